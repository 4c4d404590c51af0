/* Minimal declarations of the JNI types and JNIEnv methods that the JNI bridge
 * uses, for a compile-only check of aerodynamics4mc-fabric_amd/jni where no JDK
 * is installed. Test infrastructure: never linked into the library. */
#pragma once
#include <cstdint>
typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef int16_t jshort;
typedef float jfloat;
typedef uint8_t jboolean;
typedef jint jsize;
class _jobject {};
typedef _jobject* jobject;
typedef jobject jclass;
typedef jobject jstring;
typedef jobject jarray;
class _jbyteArray : public _jobject {};
class _jshortArray : public _jobject {};
class _jintArray : public _jobject {};
class _jfloatArray : public _jobject {};
typedef _jbyteArray* jbyteArray;
typedef _jshortArray* jshortArray;
typedef _jintArray* jintArray;
typedef _jfloatArray* jfloatArray;
#define JNI_FALSE 0
#define JNI_TRUE 1
#define JNI_ABORT 2
#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL
struct JNIEnv {
    jsize GetArrayLength(jobject);
    jbyte* GetByteArrayElements(jbyteArray, jboolean*);
    jshort* GetShortArrayElements(jshortArray, jboolean*);
    jint* GetIntArrayElements(jintArray, jboolean*);
    jfloat* GetFloatArrayElements(jfloatArray, jboolean*);
    void ReleaseByteArrayElements(jbyteArray, jbyte*, jint);
    void ReleaseShortArrayElements(jshortArray, jshort*, jint);
    void ReleaseIntArrayElements(jintArray, jint*, jint);
    void ReleaseFloatArrayElements(jfloatArray, jfloat*, jint);
    jintArray NewIntArray(jsize);
    jfloatArray NewFloatArray(jsize);
    void SetIntArrayRegion(jintArray, jsize, jsize, const jint*);
    void SetFloatArrayRegion(jfloatArray, jsize, jsize, const jfloat*);
    jstring NewStringUTF(const char*);
};
