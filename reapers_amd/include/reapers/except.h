// except.h -- exception classes of the REAPERS API (names and messages as in
// the reference, inc/except.h:27-170) plus the mapping from the C ABI's
// status codes (include/reapers_hip.h) to them.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

#define ThrowException(Ex, ...) \
    throw REAPERS::Ex(__FILE__, __LINE__, __func__, ##__VA_ARGS__)

class Exception : public std::exception {
protected:
    std::string msg;
public:
    Exception(const char *file, int line, const char *func) {
        std::stringstream ss;
        ss << file << ":" << line << "@" << func << "(): ";
        msg = ss.str();
    }
    const char *what() const noexcept override { return msg.c_str(); }
};

class IndexTooLarge : public Exception {
public:
    IndexTooLarge(const char *f, int l, const char *fn, size_t n, size_t nmax, const char *ty)
        : Exception(f, l, fn) {
        std::stringstream ss;
        ss << "Specified " << ty << " index " << n << " exceeds " << nmax << ".";
        msg += ss.str();
    }
};

class IndexTooSmall : public Exception {
public:
    IndexTooSmall(const char *f, int l, const char *fn, size_t n, size_t nmin, const char *ty)
        : Exception(f, l, fn) {
        std::stringstream ss;
        ss << "Specified " << ty << " index " << n << " cannot be less than " << nmin << ".";
        msg += ss.str();
    }
};

class SpinIndexTooLarge : public IndexTooLarge {
public:
    SpinIndexTooLarge(const char *f, int l, const char *fn, size_t n, size_t nmax)
        : IndexTooLarge(f, l, fn, n, nmax, "spin") {}
};

class FieldIndexTooLarge : public IndexTooLarge {
public:
    FieldIndexTooLarge(const char *f, int l, const char *fn, size_t n, size_t nmax)
        : IndexTooLarge(f, l, fn, n, nmax, "field") {}
};

class StateIndexTooLarge : public IndexTooLarge {
public:
    StateIndexTooLarge(const char *f, int l, const char *fn, size_t n, size_t nmax)
        : IndexTooLarge(f, l, fn, n, nmax, "state") {}
};

class SpinIndexTooSmall : public IndexTooSmall {
public:
    SpinIndexTooSmall(const char *f, int l, const char *fn, size_t n, size_t nmin)
        : IndexTooSmall(f, l, fn, n, nmin, "spin") {}
};

class FieldIndexTooSmall : public IndexTooSmall {
public:
    FieldIndexTooSmall(const char *f, int l, const char *fn, size_t n, size_t nmin)
        : IndexTooSmall(f, l, fn, n, nmin, "field") {}
};

class InvalidArgument : public Exception {
public:
    InvalidArgument(const char *f, int l, const char *fn, const char *name,
                    const char *param_msg) : Exception(f, l, fn) {
        std::stringstream ss;
        ss << "Invalid argument for " << name << ": " << name << " " << param_msg << ".";
        msg += ss.str();
    }
    InvalidArgument(const char *f, int l, const char *fn, const char *err_msg)
        : Exception(f, l, fn) {
        msg += std::string("Invalid argument: ") + err_msg;
    }
};

class RuntimeError : public Exception {
public:
    RuntimeError(const char *f, int l, const char *fn, const char *lib, int code)
        : Exception(f, l, fn) {
        std::stringstream ss;
        ss << lib << " runtime error, code " << code << ".";
        msg += ss.str();
    }
    RuntimeError(const char *f, int l, const char *fn, const std::string &detail)
        : Exception(f, l, fn) {
        msg += detail;
    }
};

// GPU runtime error reported by the HIP backend (the reference's CudaError)
class GpuError : public Exception {
public:
    GpuError(const char *f, int l, const char *fn, int code, const char *detail)
        : Exception(f, l, fn) {
        std::stringstream ss;
        ss << " GPU runtime error, code " << code << " (" << detail << ").";
        msg += ss.str();
    }
};

class NoGpuDevice : public Exception {
public:
    NoGpuDevice(const char *f, int l, const char *fn, const char *detail = "")
        : Exception(f, l, fn) {
        msg += "No GPU device available";
        if (detail && *detail) { msg += ": "; msg += detail; }
    }
};

// peer access between two devices is unavailable (the reference's
// multi-GPU split; this backend runs on one device and never throws it)
class NoP2PAccess : public Exception {
public:
    NoP2PAccess(const char *f, int l, const char *fn, int dev0, int dev1) : Exception(f, l, fn) {
        std::stringstream ss;
        ss << "P2P access between device " << dev0 << " and device " << dev1
           << " is not supported.";
        msg += ss.str();
    }
};

class DevOutOfMem : public Exception {
public:
    DevOutOfMem(const char *f, int l, const char *fn, size_t size) : Exception(f, l, fn) {
        std::stringstream ss;
        ss << "Failed to allocate " << size << " bytes on GPU device.";
        msg += ss.str();
    }
};

namespace internal {
// Throw the exception that corresponds to a C ABI status code.
[[noreturn]] inline void throw_status(const char *f, int l, const char *fn, int status,
                                      size_t bytes = 0)
{
    const char *detail = rp_last_error();
    switch (status) {
    case RP_ERR_INVALID_ARGUMENT:
        throw InvalidArgument(f, l, fn, detail);
    case RP_ERR_OUT_OF_MEMORY:
        throw DevOutOfMem(f, l, fn, bytes);
    case RP_ERR_NO_DEVICE:
        throw NoGpuDevice(f, l, fn, detail);
    default:
        throw GpuError(f, l, fn, status, detail);
    }
}
}  // namespace internal

#define RP_CALL(x)                                                              \
    do {                                                                        \
        int rp_status_ = (x);                                                   \
        if (rp_status_ != RP_OK)                                                \
            ::REAPERS::internal::throw_status(__FILE__, __LINE__, __func__, rp_status_); \
    } while (0)
