// Shared device helpers for the gfx950 kernels of the split-step propagator.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FP_DEV __device__ __forceinline__

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T> FP_DEV cpx<T> mk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <typename T> FP_DEV cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> FP_DEV cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> FP_DEV cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> FP_DEV cpx<T> operator*(T s, cpx<T> a) { return mk<T>(s * a.x, s * a.y); }
template <typename T> FP_DEV cpx<T> conjx(cpx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> FP_DEV cpx<T> mul_mi(cpx<T> a) {
  return INV ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

FP_DEV bool fp_isfinite(double v) { return __builtin_isfinite(v); }
FP_DEV bool fp_isfinite(float v) { return __builtin_isfinite(v); }

// nan_to_num(v, nan=0, posinf=0, neginf=0)
template <typename T> FP_DEV T nan0(T v) { return fp_isfinite(v) ? v : T(0); }

// Load/store helpers for interleaved complex rows.
template <typename T> FP_DEV cpx<T> ldc(const T* p, int i) {
  cpx<T> r; r.x = p[2 * i]; r.y = p[2 * i + 1]; return r;
}
template <typename T> FP_DEV void stc(T* p, int i, cpx<T> v) { p[2 * i] = v.x; p[2 * i + 1] = v.y; }
