// core.cpp -- version, build configuration and the thread-local error
// channel of the C ABI (counterpart of scl/binding/c_api/core/core.cpp:33-300).
#include "internal.hpp"

#include <cstring>
#include <new>
#include <stdexcept>

namespace scl_amd {

namespace {
constexpr size_t kMsgLen = 512;  // core.cpp:28 (ERROR_MESSAGE_BUFFER_SIZE)
thread_local scl_error_t t_code = SCL_OK;
thread_local char t_msg[kMsgLen] = {0};
}  // namespace

void set_last_error(scl_error_t code, const char* msg) noexcept {
    t_code = code;
    if (msg) {
        std::strncpy(t_msg, msg, kMsgLen - 1);
        t_msg[kMsgLen - 1] = '\0';
    } else {
        t_msg[0] = '\0';
    }
}

void clear_last_error() noexcept {
    t_code = SCL_OK;
    t_msg[0] = '\0';
}

void fail(scl_error_t code, const std::string& msg) { throw Error{code, msg}; }

scl_error_t handle_exception() noexcept {
    try {
        throw;
    } catch (const Error& e) {
        set_last_error(e.code, e.msg.c_str());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error(SCL_ERROR_OUT_OF_MEMORY, "Out of memory");
        return SCL_ERROR_OUT_OF_MEMORY;
    } catch (const std::invalid_argument& e) {
        set_last_error(SCL_ERROR_INVALID_ARGUMENT, e.what());
        return SCL_ERROR_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        set_last_error(SCL_ERROR_INDEX_OUT_OF_BOUNDS, e.what());
        return SCL_ERROR_INDEX_OUT_OF_BOUNDS;
    } catch (const std::exception& e) {
        set_last_error(SCL_ERROR_INTERNAL, e.what());
        return SCL_ERROR_INTERNAL;
    } catch (...) {
        set_last_error(SCL_ERROR_UNKNOWN, "Unknown error");
        return SCL_ERROR_UNKNOWN;
    }
}

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();  // clear the sticky-free error state
        fail(SCL_ERROR_OUT_OF_MEMORY, m);
    }
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice ||
        e == hipErrorInsufficientDriver)
        fail(SCL_ERROR_FEATURE_UNAVAILABLE, m);
    fail(SCL_ERROR_INTERNAL, m);
}

}  // namespace scl_amd

extern "C" {

SCL_EXPORT const char* scl_get_version(void) { return "1.0.0"; }

SCL_EXPORT const char* scl_get_build_config(void) {
    return SCL_REAL_TYPE_NAME "+" SCL_INDEX_TYPE_NAME "+hip-gfx950";
}

SCL_EXPORT const char* scl_get_last_error(void) {
    return scl_amd::t_msg[0] ? scl_amd::t_msg : "No error";
}

SCL_EXPORT scl_error_t scl_get_last_error_code(void) { return scl_amd::t_code; }

SCL_EXPORT void scl_clear_error(void) { scl_amd::clear_last_error(); }

SCL_EXPORT scl_bool_t scl_is_ok(scl_error_t code) {
    return code == SCL_OK ? SCL_TRUE : SCL_FALSE;
}

SCL_EXPORT scl_bool_t scl_is_error(scl_error_t code) {
    return code != SCL_OK ? SCL_TRUE : SCL_FALSE;
}

}  // extern "C"
