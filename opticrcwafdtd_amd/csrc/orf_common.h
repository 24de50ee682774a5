// Shared device helpers for the gfx950 RCWA kernels: interleaved complex128
// arithmetic (double2: x = real, y = imag), block reductions for 64-wide waves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double2 cplx;

#define ORF_DEV __device__ __forceinline__
#define ORF_HD __host__ __device__ __forceinline__

ORF_HD cplx cmk(double r, double i) { return make_double2(r, i); }
ORF_DEV cplx cadd(cplx a, cplx b) { return cmk(a.x + b.x, a.y + b.y); }
ORF_DEV cplx csub(cplx a, cplx b) { return cmk(a.x - b.x, a.y - b.y); }
ORF_DEV cplx cneg(cplx a) { return cmk(-a.x, -a.y); }
ORF_DEV cplx cconj(cplx a) { return cmk(a.x, -a.y); }
ORF_DEV cplx cmul(cplx a, cplx b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
ORF_DEV cplx cmulc(cplx a, cplx b) { return cmk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
ORF_DEV cplx cscale(cplx a, double s) { return cmk(a.x * s, a.y * s); }
ORF_DEV double cabs1(cplx a) { return fabs(a.x) + fabs(a.y); }
ORF_DEV double cabs(cplx a) { return hypot(a.x, a.y); }
ORF_DEV double cnorm2(cplx a) { return a.x * a.x + a.y * a.y; }
// a / b with Smith's scaling
ORF_DEV cplx cdiv(cplx a, cplx b) {
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return cmk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.y + b.x * r;
    return cmk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
ORF_DEV cplx crecip(cplx b) { return cdiv(cmk(1.0, 0.0), b); }
// principal square root (Re >= 0)
ORF_DEV cplx csqrt(cplx z) {
  double r = cabs(z);
  if (r == 0.0) return cmk(0.0, 0.0);
  double t = sqrt(0.5 * (r + fabs(z.x)));
  if (z.x >= 0.0) return cmk(t, z.y / (2.0 * t));
  return cmk(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}
ORF_DEV cplx cexp(cplx z) {
  double e = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return cmk(e * c, e * s);
}
// kz on the outgoing branch: Im >= 0, Re >= 0 when Im == 0 (kzBranch.m:6-8)
ORF_DEV cplx kz_branch(cplx kz2) {
  cplx kz = csqrt(kz2);
  if (kz.y < 0.0 || (kz.y == 0.0 && kz.x < 0.0)) kz = cneg(kz);
  return kz;
}
// kzNonzero.m:8-9
ORF_DEV cplx kz_nonzero(cplx kz) {
  if (cabs(kz) < 1e-8) return cmk(1e-8, 1e-8);
  return kz;
}

ORF_DEV double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
ORF_DEV double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-wide sum of a double over blockDim.x threads (<= 1024); red must hold
// blockDim.x/64 doubles.  Every thread gets the result.
ORF_DEV double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}
