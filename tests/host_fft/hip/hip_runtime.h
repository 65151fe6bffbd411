// Host stand-ins for the CPU check of the register FFT (tests/host_fft/fft_layouts.cpp):
// the transform's device functions compile as plain C++, one POSIX thread plays each
// lane, and a barrier plays the workgroup barrier.
#pragma once
#include <pthread.h>

#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline
struct float2 {
  float x, y;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }
extern pthread_barrier_t kw_host_barrier;
inline void __syncthreads() { pthread_barrier_wait(&kw_host_barrier); }
