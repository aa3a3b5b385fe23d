// Errors, device memory, the Device singleton and simulation boxes.
#include <cmath>
#include <cstdio>
#include <sstream>

#include "rumd_host.h"

void hip_check(hipError_t e, const char* cls, const char* method) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory || e == hipErrorLaunchOutOfResources || e == hipErrorInvalidConfiguration)
      throw RUMD_cudaResourceError(cls, method, hipGetErrorString(e));
    throw RUMD_Error(cls, method, std::string("HIP error: ") + hipGetErrorString(e));
  }
}

void rumdx_check(int rc, const char* cls, const char* method) {
  if (rc == RUMDX_OK) return;
  if (rc == RUMDX_E_RESOURCE) throw RUMD_cudaResourceError(cls, method, rumdx_last_error());
  throw RUMD_Error(cls, method, rumdx_last_error());
}

template <class T>
void DeviceArray<T>::resize(size_t n) {
  release();
  if (n == 0) return;
  void* p = nullptr;
  hip_check(hipMalloc(&p, n * sizeof(T)), "DeviceArray", "resize");
  hip_check(hipMemset(p, 0, n * sizeof(T)), "DeviceArray", "resize");
  // hipMemset runs on the null stream, which does not order work on
  // non-blocking streams (ReplicaBatch's): finish the fill before any kernel
  // can read the array.
  hip_check(hipStreamSynchronize(nullptr), "DeviceArray", "resize");
  ptr_ = static_cast<T*>(p);
  n_ = n;
}

template <class T>
void DeviceArray<T>::release() {
  if (ptr_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  n_ = 0;
}

template <class T>
void DeviceArray<T>::upload(const T* h, size_t n, hipStream_t st) {
  if (n > n_) throw RUMD_Error("DeviceArray", "upload", "upload larger than allocation");
  if (n == 0) return;
  hip_check(hipMemcpyAsync(ptr_, h, n * sizeof(T), hipMemcpyHostToDevice, st), "DeviceArray", "upload");
  hip_check(hipStreamSynchronize(st), "DeviceArray", "upload");
}

template <class T>
void DeviceArray<T>::download(T* h, size_t n, hipStream_t st) const {
  if (n > n_) throw RUMD_Error("DeviceArray", "download", "download larger than allocation");
  if (n == 0) return;
  hip_check(hipMemcpyAsync(h, ptr_, n * sizeof(T), hipMemcpyDeviceToHost, st), "DeviceArray", "download");
  hip_check(hipStreamSynchronize(st), "DeviceArray", "download");
}

template class DeviceArray<float>;
template class DeviceArray<double>;
template class DeviceArray<uint32_t>;
template class DeviceArray<uint16_t>;
template class DeviceArray<char>;
template class DeviceArray<rumdx_nh_state>;
template class DeviceArray<rumdx_replica>;

// ------------------------------------------------------------------ Device
Device& Device::GetDevice() {
  static Device d;
  return d;
}

void Device::Init() {
  if (initialized_) return;
  int n = 0;
  hip_check(hipGetDeviceCount(&n), "Device", "Init");
  if (n < 1) throw RUMD_Error("Device", "Init", "no HIP device visible");
  int dev = 0;
  hip_check(hipGetDevice(&dev), "Device", "Init");
  hip_check(hipGetDeviceProperties(&prop_, dev), "Device", "Init");
  hip_check(hipEventCreate(&start_), "Device", "Init");
  hip_check(hipEventRecord(start_, nullptr), "Device", "Init");
  initialized_ = true;
}

std::string Device::GetDeviceName() {
  Init();
  return prop_.name;
}

std::string Device::GetDeviceReport() {
  Init();
  std::ostringstream s;
  s << prop_.name << " (" << prop_.gcnArchName << "), " << prop_.multiProcessorCount << " CUs, "
    << (prop_.totalGlobalMem >> 20) << " MiB";
  return s.str();
}

size_t Device::GetDeviceMemory() {
  Init();
  return prop_.totalGlobalMem;
}

void Device::Synchronize() { hip_check(hipDeviceSynchronize(), "Device", "Synchronize"); }

float Device::Time() {
  Init();
  hipEvent_t now;
  hip_check(hipEventCreate(&now), "Device", "Time");
  hip_check(hipEventRecord(now, nullptr), "Device", "Time");
  hip_check(hipEventSynchronize(now), "Device", "Time");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, start_, now), "Device", "Time");
  (void)hipEventDestroy(now);
  return ms * 1e-3f;
}

void Device::CheckErrors() {
  hip_check(hipDeviceSynchronize(), "Device", "CheckErrors");
  hip_check(hipGetLastError(), "Device", "CheckErrors");
}

// ------------------------------------------------------------------ boxes
RectangularSimulationBox::RectangularSimulationBox(float Lx, float Ly, float Lz) { SetLengths(Lx, Ly, Lz); }

void RectangularSimulationBox::SetLengths(float Lx, float Ly, float Lz) {
  if (!(Lx > 0.f && Ly > 0.f && Lz > 0.f))
    throw RUMD_Error("RectangularSimulationBox", "SetLengths", "box lengths must be positive");
  box8_[1] = Lx;
  box8_[2] = Ly;
  box8_[3] = Lz;
  box8_[4] = 1.f / Lx;
  box8_[5] = 1.f / Ly;
  box8_[6] = 1.f / Lz;
  box8_[7] = 0.f;
  // cubic flag with RUMD's tolerance (src/SimulationBox.cu:31-43)
  const bool cubic = std::fabs(Lx / Ly - 1.) < 1e-8 && std::fabs(Ly / Lz - 1.) < 1e-8;
  box8_[0] = cubic ? 1.f : 0.f;
}

float RectangularSimulationBox::GetVolume() const { return box8_[1] * box8_[2] * box8_[3]; }

float RectangularSimulationBox::GetLength(int dir) const {
  if (dir < 0 || dir > 2) throw RUMD_Error("RectangularSimulationBox", "GetLength", "direction must be 0, 1 or 2");
  return box8_[1 + dir];
}

std::string RectangularSimulationBox::GetInfoString(unsigned precision) const {
  std::ostringstream s;
  s.precision(precision);
  s << "RectangularSimulationBox," << box8_[1] << "," << box8_[2] << "," << box8_[3];
  return s.str();
}

void Potential::SetSample(Sample* s) {
  sample = s;
  Initialize();
}
