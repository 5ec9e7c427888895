// hs_core.hip -- library-level state of libhsmooth.so: device selection,
// error reporting and the device-memory counter.
#include "hs_common.hpp"

#include <mutex>

namespace hs {

std::atomic<int64_t> g_device_bytes{0};

namespace {
std::atomic<uint64_t> g_serial{0};
}
uint64_t next_serial() { return ++g_serial; }

namespace {
thread_local std::string t_err;
thread_local int t_code = HS_OK;
thread_local int64_t t_aux = 0;
}  // namespace

void set_error(int code, const std::string &msg, int64_t aux) {
    t_code = code;
    t_err = msg;
    t_aux = aux;
}

int last_code() { return t_code; }

}  // namespace hs

extern "C" {

int hs_init(int device) {
    return hs::guard([&] {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        HS_REQUIRE(e == hipSuccess && count > 0, HS_ENODEVICE,
                   "ROCBackend is not functional: no HIP device is visible");
        HS_REQUIRE(device >= 0 && device < count, HS_EINVAL, "device index out of range");
        HS_CHECK_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        HS_CHECK_HIP(hipGetDeviceProperties(&prop, device));
        HS_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, HS_ENODEVICE,
                   std::string("libhsmooth is built for gfx950; device is ") + prop.gcnArchName);
        // this device's address of the allocation-free entry point's scratch
        hs::noalloc_scratch_init();
    });
}

const char *hs_last_error(void) { return hs::t_err.c_str(); }

int64_t hs_last_error_aux(void) { return hs::t_aux; }

int64_t hs_device_bytes(void) { return hs::g_device_bytes.load(); }

}  // extern "C"
