// CPU emulation of the HIP subset the FDTD library uses, for testing the kernels
// on the host: device memory is host memory, a launch walks its blocks in order,
// and each 64-lane wavefront runs on 64 host threads in lockstep wherever the
// kernels exchange values (__shfl_up/__shfl_down go through a barrier), so the
// results equal the GPU's up to floating-point contraction.  Built with
// AddressSanitizer it also checks every index the kernels form.  Graph capture is
// reported as unavailable, so batches run their steps one launch at a time.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __builtin_amdgcn_readfirstlane(x) (x)
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct uint3e { unsigned x, y, z; };
extern thread_local uint3e blockIdx, threadIdx, blockDim, gridDim;
inline float atomicAdd(float* p, float v) { float o = *p; *p += v; return o; }
// lockstep wavefront emulation: each of 64 worker threads is one lane
float emu_exchange(float v, int src);
extern thread_local int emu_lane;
inline float __shfl_up(float v, unsigned d, int w) { int l = emu_lane; return emu_exchange(v, (l % w) >= (int)d ? l - (int)d : l); }
inline float __shfl_down(float v, unsigned d, int w) { int l = emu_lane; return emu_exchange(v, (l % w) + (int)d < w ? l + (int)d : l); }
using std::fmax; using std::fmin; using std::exp; using std::cos;
inline float fminf(float a, float b) { return a < b ? a : b; }
template <typename T> inline T min(T a, T b) { return a < b ? a : b; }
template <typename T> inline T max(T a, T b) { return a > b ? a : b; }
inline float fmaxf(float a, float b) { return a > b ? a : b; }
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum { hipSuccess = 0, hipErrorUnknown = 999 };
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2, hipStreamCaptureModeRelaxed = 2 };
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipMalloc(void** p, size_t n) { *p = std::malloc(n ? n : 1); return *p ? hipSuccess : hipErrorUnknown; }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, int) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (void*)1; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int, hipStream_t) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
typedef int hipDevice_t;
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
inline hipError_t hipStreamGetDevice(hipStream_t, hipDevice_t* d) { *d = 0; return hipSuccess; }
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* st) { *st = hipStreamCaptureStatusNone; return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) { *p = const_cast<void*>(symbol); return hipSuccess; }
#define HIP_SYMBOL(x) (static_cast<const void*>(&(x)))
#define __align__(n) __attribute__((aligned(n)))
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipErrorUnknown; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorUnknown; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, int) { return hipErrorUnknown; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorUnknown; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
#include <functional>
void emu_run_grid(dim3 g, dim3 b, const std::function<void()>& job);
template <typename F, typename... A>
void emu_launch(F f, dim3 g, dim3 b, A... args) {
  emu_run_grid(g, b, [=]() { f(args...); });
}
// every launch is counted by kernel name (emu_kernel_launches), so a test can
// tell which kernels a case ran
void emu_note_launch(const char* name);
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) (emu_note_launch(#K), emu_launch(K, dim3(G), dim3(B), __VA_ARGS__))
