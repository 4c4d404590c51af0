// Test infrastructure (the checker, not the product): C entry points over the
// reference's analysis codec (native/src/aero_lbm_analysis_codec.cpp), which this
// file is linked with, together with the zfp 1.0.1 sources the reference vendors
// (native/third_party/zfp), into oracle/_ref/libcodec_ref.so. Nothing here is
// shipped or used on the product path.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "aero_lbm_analysis_codec.hpp"
#include "zfp.h"

extern "C" {

// returns the compressed size in bytes (0 on failure); writes at most `capacity` bytes
long long ref_codec_compress_3f(const float* values, int nx, int ny, int nz, double tolerance, uint8_t* out,
                                long long capacity) {
    std::vector<std::uint8_t> bytes;
    std::string error;
    if (!aero_lbm_analysis_codec::compress_float_grid_3d(values, nx, ny, nz, tolerance, bytes, error)) return 0;
    if ((long long)bytes.size() > capacity) return -(long long)bytes.size();
    std::memcpy(out, bytes.data(), bytes.size());
    return (long long)bytes.size();
}

// zfp's own API in another mode (0 fixed rate, 1 fixed precision, 2 fixed accuracy),
// full header, for decoder coverage beyond the reference's fixed-accuracy streams
long long ref_zfp_compress_3f_mode(const float* values, int nx, int ny, int nz, int mode, double param, uint8_t* out,
                                   long long capacity) {
    zfp_field* field = zfp_field_3d(const_cast<float*>(values), zfp_type_float, (size_t)nx, (size_t)ny, (size_t)nz);
    zfp_stream* zs = zfp_stream_open(nullptr);
    if (mode == 0) zfp_stream_set_rate(zs, param, zfp_type_float, 3, zfp_false);
    else if (mode == 1) zfp_stream_set_precision(zs, (unsigned)param);
    else zfp_stream_set_accuracy(zs, param);
    const size_t max = zfp_stream_maximum_size(zs, field);
    long long n = 0;
    if (max && (long long)max <= capacity) {
        std::memset(out, 0, (size_t)capacity);
        bitstream* bs = stream_open(out, (size_t)capacity);
        zfp_stream_set_bit_stream(zs, bs);
        zfp_stream_rewind(zs);
        if (zfp_write_header(zs, field, ZFP_HEADER_FULL)) n = (long long)zfp_compress(zs, field);
        stream_close(bs);
    }
    zfp_stream_close(zs);
    zfp_field_free(field);
    return n;
}

int ref_codec_decompress_3f(const uint8_t* data, long long size, int nx, int ny, int nz, float* out) {
    std::string error;
    return aero_lbm_analysis_codec::decompress_float_grid_3d(data, (std::size_t)size, nx, ny, nz, out, error) ? 1 : 0;
}

}
