/* JNI types and the JNIEnv interface function table, written from the Java Native
 * Interface Specification ("JNI Types and Data Structures", "JNI Functions": each
 * function's LINKAGE gives its index in the JNIEnv interface function table). A JVM
 * hands native methods a JNIEnv* whose first word points at that table; this header
 * names the slots the bridge calls and leaves the others as opaque pointers, so the
 * bridge compiles without a JDK and binds to any JVM's table by index. */
#ifndef AERO_LBM_JNI_H
#define AERO_LBM_JNI_H

#include <stddef.h>

#include <stdarg.h>
#include <stdint.h>

#define JNIEXPORT __attribute__((visibility("default")))
#define JNIIMPORT
#define JNICALL

typedef unsigned char jboolean;
typedef signed char jbyte;
typedef unsigned short jchar;
typedef short jshort;
typedef int32_t jint;
typedef int64_t jlong;
typedef float jfloat;
typedef double jdouble;
typedef jint jsize;

#define JNI_FALSE 0
#define JNI_TRUE 1
/* release modes of Release<Type>ArrayElements */
#define JNI_COMMIT 1
#define JNI_ABORT 2
#define JNI_OK 0
#define JNI_VERSION_1_8 0x00010008

#ifdef __cplusplus
class _jobject {};
class _jclass : public _jobject {};
class _jstring : public _jobject {};
class _jarray : public _jobject {};
class _jbooleanArray : public _jarray {};
class _jbyteArray : public _jarray {};
class _jcharArray : public _jarray {};
class _jshortArray : public _jarray {};
class _jintArray : public _jarray {};
class _jlongArray : public _jarray {};
class _jfloatArray : public _jarray {};
class _jdoubleArray : public _jarray {};
typedef _jobject* jobject;
typedef _jclass* jclass;
typedef _jstring* jstring;
typedef _jarray* jarray;
typedef _jbooleanArray* jbooleanArray;
typedef _jbyteArray* jbyteArray;
typedef _jcharArray* jcharArray;
typedef _jshortArray* jshortArray;
typedef _jintArray* jintArray;
typedef _jlongArray* jlongArray;
typedef _jfloatArray* jfloatArray;
typedef _jdoubleArray* jdoubleArray;
#else
struct _jobject;
typedef struct _jobject* jobject;
typedef jobject jclass;
typedef jobject jstring;
typedef jobject jarray;
typedef jarray jbooleanArray;
typedef jarray jbyteArray;
typedef jarray jcharArray;
typedef jarray jshortArray;
typedef jarray jintArray;
typedef jarray jlongArray;
typedef jarray jfloatArray;
typedef jarray jdoubleArray;
#endif

struct JNINativeInterface_;
#ifdef __cplusplus
struct JNIEnv_;
typedef JNIEnv_ JNIEnv;
#else
typedef const struct JNINativeInterface_* JNIEnv;
#endif

/* The interface function table. Slots the bridge does not call are opaque; the
 * comments give the specification's index of the first slot of each run. */
struct JNINativeInterface_ {
    void* slots_0_166[167]; /* 0-3 reserved, 4 GetVersion ... 166 ReleaseStringChars */
    jstring(JNICALL* NewStringUTF)(JNIEnv* env, const char* utf);                   /* 167 */
    void* slots_168_170[3];                                                          /* 168 GetStringUTFLength */
    jsize(JNICALL* GetArrayLength)(JNIEnv* env, jarray array);                       /* 171 */
    void* slots_172_175[4];                                                          /* 172 NewObjectArray */
    jbyteArray(JNICALL* NewByteArray)(JNIEnv* env, jsize len);                       /* 176 */
    void* slots_177_178[2];                                                          /* 177 NewCharArray */
    jintArray(JNICALL* NewIntArray)(JNIEnv* env, jsize len);                         /* 179 */
    void* slot_180;                                                                  /* 180 NewLongArray */
    jfloatArray(JNICALL* NewFloatArray)(JNIEnv* env, jsize len);                     /* 181 */
    void* slots_182_183[2];                                                          /* 182 NewDoubleArray */
    jbyte*(JNICALL* GetByteArrayElements)(JNIEnv* env, jbyteArray a, jboolean* c);   /* 184 */
    void* slot_185;                                                                  /* 185 GetCharArrayElements */
    jshort*(JNICALL* GetShortArrayElements)(JNIEnv* env, jshortArray a, jboolean* c); /* 186 */
    jint*(JNICALL* GetIntArrayElements)(JNIEnv* env, jintArray a, jboolean* c);      /* 187 */
    void* slot_188;                                                                  /* 188 GetLongArrayElements */
    jfloat*(JNICALL* GetFloatArrayElements)(JNIEnv* env, jfloatArray a, jboolean* c); /* 189 */
    void* slots_190_191[2];                                                          /* 190 GetDoubleArrayElements */
    void(JNICALL* ReleaseByteArrayElements)(JNIEnv* env, jbyteArray a, jbyte* e, jint mode);   /* 192 */
    void* slot_193;                                                                             /* 193 Char */
    void(JNICALL* ReleaseShortArrayElements)(JNIEnv* env, jshortArray a, jshort* e, jint mode); /* 194 */
    void(JNICALL* ReleaseIntArrayElements)(JNIEnv* env, jintArray a, jint* e, jint mode);       /* 195 */
    void* slot_196;                                                                             /* 196 Long */
    void(JNICALL* ReleaseFloatArrayElements)(JNIEnv* env, jfloatArray a, jfloat* e, jint mode); /* 197 */
    void* slots_198_199[2];                                                          /* 198 ReleaseDouble... */
    void(JNICALL* GetByteArrayRegion)(JNIEnv* env, jbyteArray a, jsize start, jsize len, jbyte* buf);       /* 200 */
    void* slots_201_207[7];                                                                                 /* 201 */
    void(JNICALL* SetByteArrayRegion)(JNIEnv* env, jbyteArray a, jsize start, jsize len, const jbyte* buf); /* 208 */
    void* slots_209_210[2];                                                                                 /* 209 */
    void(JNICALL* SetIntArrayRegion)(JNIEnv* env, jintArray a, jsize start, jsize len, const jint* buf);     /* 211 */
    void* slot_212;                                                                                         /* 212 */
    void(JNICALL* SetFloatArrayRegion)(JNIEnv* env, jfloatArray a, jsize start, jsize len, const jfloat* buf); /* 213 */
    void* slots_214_227[14]; /* 214 SetDoubleArrayRegion ... 227 DeleteWeakGlobalRef */
    jboolean(JNICALL* ExceptionCheck)(JNIEnv* env); /* 228 */
    void* slots_229_234[6];  /* 229 NewDirectByteBuffer ... 234 IsVirtualThread */
};

#ifdef __cplusplus
/* slot positions the bridge relies on */
static_assert(offsetof(JNINativeInterface_, NewStringUTF) == 167 * sizeof(void*), "JNI table layout");
static_assert(offsetof(JNINativeInterface_, NewByteArray) == 176 * sizeof(void*), "JNI table layout");
static_assert(offsetof(JNINativeInterface_, GetByteArrayRegion) == 200 * sizeof(void*), "JNI table layout");
static_assert(offsetof(JNINativeInterface_, SetByteArrayRegion) == 208 * sizeof(void*), "JNI table layout");
static_assert(offsetof(JNINativeInterface_, SetFloatArrayRegion) == 213 * sizeof(void*), "JNI table layout");
static_assert(offsetof(JNINativeInterface_, ExceptionCheck) == 228 * sizeof(void*), "JNI table layout");
static_assert(sizeof(JNINativeInterface_) == 235 * sizeof(void*), "JNI table layout");
#endif

#ifdef __cplusplus
/* C++ view of JNIEnv: the table pointer plus inline forwarding members, as the
 * specification's C++ binding provides */
struct JNIEnv_ {
    const JNINativeInterface_* functions;
    jstring NewStringUTF(const char* s) { return functions->NewStringUTF(this, s); }
    jsize GetArrayLength(jarray a) { return functions->GetArrayLength(this, a); }
    jintArray NewIntArray(jsize n) { return functions->NewIntArray(this, n); }
    jfloatArray NewFloatArray(jsize n) { return functions->NewFloatArray(this, n); }
    jbyteArray NewByteArray(jsize n) { return functions->NewByteArray(this, n); }
    void GetByteArrayRegion(jbyteArray a, jsize s, jsize n, jbyte* b) { functions->GetByteArrayRegion(this, a, s, n, b); }
    void SetByteArrayRegion(jbyteArray a, jsize s, jsize n, const jbyte* b) {
        functions->SetByteArrayRegion(this, a, s, n, b);
    }
    jboolean ExceptionCheck() { return functions->ExceptionCheck(this); }
    jbyte* GetByteArrayElements(jbyteArray a, jboolean* c) { return functions->GetByteArrayElements(this, a, c); }
    jshort* GetShortArrayElements(jshortArray a, jboolean* c) { return functions->GetShortArrayElements(this, a, c); }
    jint* GetIntArrayElements(jintArray a, jboolean* c) { return functions->GetIntArrayElements(this, a, c); }
    jfloat* GetFloatArrayElements(jfloatArray a, jboolean* c) { return functions->GetFloatArrayElements(this, a, c); }
    void ReleaseByteArrayElements(jbyteArray a, jbyte* e, jint m) { functions->ReleaseByteArrayElements(this, a, e, m); }
    void ReleaseShortArrayElements(jshortArray a, jshort* e, jint m) {
        functions->ReleaseShortArrayElements(this, a, e, m);
    }
    void ReleaseIntArrayElements(jintArray a, jint* e, jint m) { functions->ReleaseIntArrayElements(this, a, e, m); }
    void ReleaseFloatArrayElements(jfloatArray a, jfloat* e, jint m) {
        functions->ReleaseFloatArrayElements(this, a, e, m);
    }
    void SetIntArrayRegion(jintArray a, jsize s, jsize n, const jint* b) { functions->SetIntArrayRegion(this, a, s, n, b); }
    void SetFloatArrayRegion(jfloatArray a, jsize s, jsize n, const jfloat* b) {
        functions->SetFloatArrayRegion(this, a, s, n, b);
    }
};
#endif

#endif /* AERO_LBM_JNI_H */
