// Core device types of the MI355X (gfx950) matrix-element kernels.
// One event per lane: every wavefunction is a handful of complex values kept in
// VGPRs; nothing intermediate is written to HBM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfx {

template <class R>
struct cx {
  R re, im;
  __device__ __forceinline__ cx() {}
  __device__ __forceinline__ constexpr cx(R r, R i) : re(r), im(i) {}
};

template <class R> __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return cx<R>(a.re + b.re, a.im + b.im); }
template <class R> __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return cx<R>(a.re - b.re, a.im - b.im); }
template <class R> __device__ __forceinline__ cx<R> operator-(cx<R> a) { return cx<R>(-a.re, -a.im); }
template <class R> __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class R> __device__ __forceinline__ cx<R> operator*(cx<R> a, R s) { return cx<R>(a.re * s, a.im * s); }
template <class R> __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return cx<R>(a.re * s, a.im * s); }
template <class R> __device__ __forceinline__ cx<R>& operator+=(cx<R>& a, cx<R> b) { a.re += b.re; a.im += b.im; return a; }
template <class R> __device__ __forceinline__ cx<R>& operator-=(cx<R>& a, cx<R> b) { a.re -= b.re; a.im -= b.im; return a; }
// multiplication by +i / -i: a component swap
template <class R> __device__ __forceinline__ cx<R> mul_i(cx<R> a) { return cx<R>(-a.im, a.re); }
template <class R> __device__ __forceinline__ cx<R> mul_mi(cx<R> a) { return cx<R>(a.im, -a.re); }
template <class R> __device__ __forceinline__ R recip(R x) { return R(1) / x; }
// FP32: the hardware reciprocal (v_rcp_f32, within 1 ulp) instead of the IEEE division
// sequence; the FP64 build keeps the correctly rounded division of the reference.
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <class R> __device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) {
  R inv = recip(b.re * b.re + b.im * b.im);
  return cx<R>((a.re * b.re + a.im * b.im) * inv, (a.im * b.re - a.re * b.im) * inv);
}

// A HELAS wavefunction: slots [0],[1] carry the momentum, [2..5] the spinor/vector.
template <class R>
struct W {
  cx<R> v[6];
  __device__ __forceinline__ cx<R>& operator[](int i) { return v[i]; }
  __device__ __forceinline__ const cx<R>& operator[](int i) const { return v[i]; }
};

// Model parameters in the order the emitted modules pass them (mdl_MT, mdl_WT,
// GC_10, GC_11, GC_12), plus the value used for sqrt(1/2) in vxxxxx.
template <class R>
struct Params {
  R mdl_MT, mdl_WT;
  cx<R> GC_10, GC_11, GC_12;
  R sqh;
};

// Arithmetic forms per precision.  FP32 (value = true) uses cancellation-free forms:
// propagator P^2 - M^2 from per-event invariants 2 k_i.k_j computed in FP64 (the m_t^2
// terms cancel exactly, no cancellation in t-channel denominators) and |p| + p_z of
// massive spinors as pT^2 / (|p| - p_z) for p_z < 0.  FP64 (value = false) uses the
// reference's own expressions (Lorentz square of the current's slot momentum, |p| + p_z),
// so the FP64 build reproduces the reference to rounding.
template <class R> struct StableForms { static constexpr bool value = false; };
template <> struct StableForms<float> { static constexpr bool value = true; };

}  // namespace mfx
