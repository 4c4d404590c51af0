/* D3Q19 mesoscale oracle -- TEST INFRASTRUCTURE ONLY (filled in with the mesoscale path). */
int meso_oracle_version(void) { return 1; }
