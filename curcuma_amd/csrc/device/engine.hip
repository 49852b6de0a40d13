// placeholder, replaced by the kernel implementation
#include <hip/hip_runtime.h>
#include "engine.hpp"
namespace gfnff {
struct DeviceEngine { int dummy; };
DeviceEngine* device_create(const ParamSet&, int, int& status, std::string& msg) { status = 5; msg = "not built"; return nullptr; }
void device_destroy(DeviceEngine* e) { delete e; }
int device_eval_host(DeviceEngine*, const double*, bool, double*, double*, double*, double*, std::string& msg) { msg = "not built"; return 5; }
int device_eval(DeviceEngine*, const double*, double*, double*, void*, std::string& msg) { msg = "not built"; return 5; }
int device_count() { int n = 0; if (hipGetDeviceCount(&n) != hipSuccess) return 0; return n; }
}
