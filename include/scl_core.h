/*
 * scl_core.h -- core types, error codes and error channel of the scl C ABI,
 * as shipped by this library (the f32_i64 variant: scl_real_t = float,
 * scl_index_t = int64_t).
 *
 * Replaces: scl/binding/c_api/core/core.h:47-187 (export macro, version,
 * opaque handles, scl_real_t / scl_index_t / scl_size_t / scl_bool_t,
 * scl_error_t and the numeric error codes) and core.cpp:33-75, 267-300
 * (thread-local last error, version and build-config strings).
 * Codes and typedefs are identical; only the f32_i64 variant is provided.
 */
#ifndef SCL_CORE_H
#define SCL_CORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define SCL_EXPORT __attribute__((visibility("default")))
#else
#define SCL_EXPORT
#endif

#define SCL_C_API_VERSION_MAJOR 1
#define SCL_C_API_VERSION_MINOR 0
#define SCL_C_API_VERSION_PATCH 0

/* core.h:67-73 */
typedef struct scl_sparse_matrix scl_sparse_matrix;
typedef scl_sparse_matrix* scl_sparse_t;

/* core.h:81-111 with SCL_USE_FLOAT32 + SCL_USE_INT64 */
typedef float scl_real_t;
#define SCL_REAL_TYPE_NAME "float32"
typedef int64_t scl_index_t;
#define SCL_INDEX_TYPE_NAME "int64"
typedef size_t scl_size_t;
typedef int scl_bool_t;
#define SCL_TRUE 1
#define SCL_FALSE 0

/* core.h:126-166 */
typedef int32_t scl_error_t;
#define SCL_OK 0
#define SCL_ERROR_UNKNOWN 1
#define SCL_ERROR_INTERNAL 2
#define SCL_ERROR_OUT_OF_MEMORY 3
#define SCL_ERROR_NULL_POINTER 4
#define SCL_ERROR_INVALID_ARGUMENT 10
#define SCL_ERROR_DIMENSION_MISMATCH 11
#define SCL_ERROR_DOMAIN_ERROR 12
#define SCL_ERROR_RANGE_ERROR 13
#define SCL_ERROR_INDEX_OUT_OF_BOUNDS 14
#define SCL_ERROR_TYPE_ERROR 20
#define SCL_ERROR_TYPE_MISMATCH 21
#define SCL_ERROR_IO_ERROR 30
#define SCL_ERROR_NOT_IMPLEMENTED 40
#define SCL_ERROR_FEATURE_UNAVAILABLE 41
#define SCL_ERROR_NUMERICAL_ERROR 50
#define SCL_ERROR_DIVISION_BY_ZERO 51
#define SCL_ERROR_OVERFLOW 52
#define SCL_ERROR_UNDERFLOW 53
#define SCL_ERROR_CONVERGENCE_ERROR 54

/* core.cpp:267-270 */
SCL_EXPORT const char* scl_get_version(void);
/* core.cpp:272-300; here e.g. "float32+int64+hip-gfx950" */
SCL_EXPORT const char* scl_get_build_config(void);

/* core.h: thread-local error channel (core.cpp:33-75) */
SCL_EXPORT const char* scl_get_last_error(void);
SCL_EXPORT scl_error_t scl_get_last_error_code(void);
SCL_EXPORT void scl_clear_error(void);
SCL_EXPORT scl_bool_t scl_is_ok(scl_error_t code);
SCL_EXPORT scl_bool_t scl_is_error(scl_error_t code);

#ifdef __cplusplus
}
#endif
#endif
