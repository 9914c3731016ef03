// fgpe_device.hpp -- device building blocks of libfgpe for gfx950 (CDNA4).
//
//  * complex arithmetic on float2/double2 (interleaved re/im, 8/16-byte loads)
//  * in-register small DFTs (radix 2, 3, 4, 6, 8, 12, 16)
//  * a Stockham line FFT in which each of the P = N/Q threads of a line keeps
//    Q points in registers and exchanges through LDS between stages
//  * the pointwise physics of one grid point (SE, SP, SOC) in registers
//
// The FFT replaces FFTW's fftw_execute_dft (SRC/cgpe3d.f90:1760,1781): an
// unnormalised DFT, forward sign -1, backward +1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fgpe {

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).  Used
// wherever a register array is indexed, so the index is a constant and the
// array stays in VGPRs (a rolled loop would demote it to scratch memory).
template <int B, int E> struct SFor {
  template <typename F> __host__ __device__ __forceinline__ static void run(F&& f) {
    if constexpr (B < E) {
      f(std::integral_constant<int, B>{});
      SFor<B + 1, E>::run(f);
    }
  }
};

template <typename T> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename T> using C_t = typename cplx_of<T>::type;

template <typename C> __host__ __device__ __forceinline__ C mk(decltype(C::x) re, decltype(C::x) im) {
  C c;
  c.x = re;
  c.y = im;
  return c;
}
template <typename C> __host__ __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __host__ __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __host__ __device__ __forceinline__ C cmul(C a, C b) {
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename C> __host__ __device__ __forceinline__ C cmulc(C a, C b) {
  return mk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename C> __host__ __device__ __forceinline__ C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C, typename T> __host__ __device__ __forceinline__ C cscale(C a, T s) { return mk<C>(a.x * s, a.y * s); }
// i*a and -i*a
template <typename C> __host__ __device__ __forceinline__ C cmuli(C a) { return mk<C>(-a.y, a.x); }
template <typename C> __host__ __device__ __forceinline__ C cmulmi(C a) { return mk<C>(a.y, -a.x); }
template <typename C> __host__ __device__ __forceinline__ decltype(C::x) cabs2(C a) { return a.x * a.x + a.y * a.y; }

// ------------------------------------------------- unbiased constant rotations
// A constant factor w of every step (FFT twiddles, kinetic factors) stored,
// in FP32, as hi + lo parts that give its FP64 value.  The product is formed
// by an FMA chain that folds the lo parts in before the last rounding, so the
// result is rounded once from (nearly) the exact product.  A plain FP32 factor
// would carry the same rounding of w into every step (|fl(w)| != 1), a bias
// that accumulates linearly over the run, where unbiased rounding only
// random-walks.  FP64 keeps a single part.
template <typename T> struct rot_of;
template <> struct rot_of<float> { using type = float4; };   // re_hi, im_hi, re_lo, im_lo
template <> struct rot_of<double> { using type = double2; };  // re, im
template <typename T> using R_t = typename rot_of<T>::type;

// v * w (DIR < 0) or v * conj(w) (DIR > 0)
template <int DIR> __host__ __device__ __forceinline__ float2 rot(float2 v, float4 w) {
  const float wi = DIR > 0 ? -w.y : w.y, wil = DIR > 0 ? -w.w : w.w;
  const float re = fmaf(v.x, w.x, fmaf(-v.y, wi, fmaf(v.x, w.z, -v.y * wil)));
  const float im = fmaf(v.x, wi, fmaf(v.y, w.x, fmaf(v.x, wil, v.y * w.z)));
  return make_float2(re, im);
}
template <int DIR> __host__ __device__ __forceinline__ double2 rot(double2 v, double2 w) {
  return DIR > 0 ? cmulc(v, w) : cmul(v, w);
}
// v * (1 + h) = v + v h for a unit-modulus factor near 1 stored as h = w - 1
// (hi + lo parts in FP32).  The change v h is formed first and added last, so
// the one rounding at full magnitude happens on a value that varies from step
// to step.  Applying w itself would round v * fl(cos) with fl(cos) == 1 for
// small angles: the product is then exact, the lo parts fall below half an ulp
// and are dropped the same way every step (a norm drift of order angle^2 per
// step for the slowly rotating low modes that carry the condensate).
__host__ __device__ __forceinline__ float2 rot1(float2 v, float4 h) {
  const float dr = fmaf(v.x, h.x, fmaf(-v.y, h.y, fmaf(v.x, h.z, -v.y * h.w)));
  const float di = fmaf(v.x, h.y, fmaf(v.y, h.x, fmaf(v.x, h.w, v.y * h.z)));
  return make_float2(v.x + dr, v.y + di);
}
__host__ __device__ __forceinline__ double2 rot1(double2 v, double2 h) { return cadd(v, cmul(v, h)); }

// v * (c + i s) for FP64 constants c, s (compile-time after unrolling)
template <typename C> __host__ __device__ __forceinline__ C rotc(C v, double c, double s) {
  using T = decltype(C::x);
  if constexpr (sizeof(T) == 4) {
    const float ch = (float)c, cl = (float)(c - (double)ch), sh = (float)s, sl = (float)(s - (double)sh);
    const float re = fmaf(v.x, ch, fmaf(-v.y, sh, fmaf(v.x, cl, -v.y * sl)));
    const float im = fmaf(v.x, sh, fmaf(v.y, ch, fmaf(v.x, sl, v.y * cl)));
    return mk<C>(re, im);
  } else {
    return mk<C>(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}
// v * a for a real FP64 constant a
template <typename C> __host__ __device__ __forceinline__ C scalec(C v, double a) {
  using T = decltype(C::x);
  if constexpr (sizeof(T) == 4) {
    const float h = (float)a, l = (float)(a - (double)h);
    return mk<C>(fmaf(v.x, h, v.x * l), fmaf(v.y, h, v.y * l));
  } else {
    return mk<C>(v.x * a, v.y * a);
  }
}

// ---------------------------------------------------------------- small DFTs
// Exact twiddle constants omega_R^e = cos(2 pi e/R) + DIR i sin(2 pi e/R).
__host__ __device__ constexpr double cos_tab(int e, int R) {
  // first-octant values, extended by symmetry
  constexpr double C16_1 = 0.92387953251128675613, S16_1 = 0.38268343236508977173;
  constexpr double H = 0.70710678118654752440, R3 = 0.86602540378443864676;
  e = ((e % R) + R) % R;
  if (R == 16) {
    constexpr double t[16] = {1, C16_1, H, S16_1, 0, -S16_1, -H, -C16_1, -1, -C16_1, -H, -S16_1, 0, S16_1, H, C16_1};
    return t[e];
  }
  if (R == 12) {
    constexpr double t[12] = {1, R3, 0.5, 0, -0.5, -R3, -1, -R3, -0.5, 0, 0.5, R3};
    return t[e];
  }
  if (R == 8) {
    constexpr double t[8] = {1, H, 0, -H, -1, -H, 0, H};
    return t[e];
  }
  if (R == 6) {
    constexpr double t[6] = {1, 0.5, -0.5, -1, -0.5, 0.5};
    return t[e];
  }
  if (R == 4) {
    constexpr double t[4] = {1, 0, -1, 0};
    return t[e];
  }
  if (R == 3) {
    constexpr double t[3] = {1, -0.5, -0.5};
    return t[e];
  }
  constexpr double t2[2] = {1, -1};
  return t2[e];
}
__host__ __device__ constexpr double sin_tab(int e, int R) {
  constexpr double C16_1 = 0.92387953251128675613, S16_1 = 0.38268343236508977173;
  constexpr double H = 0.70710678118654752440, R3 = 0.86602540378443864676;
  e = ((e % R) + R) % R;
  if (R == 16) {
    constexpr double t[16] = {0, S16_1, H, C16_1, 1, C16_1, H, S16_1, 0, -S16_1, -H, -C16_1, -1, -C16_1, -H, -S16_1};
    return t[e];
  }
  if (R == 12) {
    constexpr double t[12] = {0, 0.5, R3, 1, R3, 0.5, 0, -0.5, -R3, -1, -R3, -0.5};
    return t[e];
  }
  if (R == 8) {
    constexpr double t[8] = {0, H, 1, H, 0, -H, -1, -H};
    return t[e];
  }
  if (R == 6) {
    constexpr double t[6] = {0, R3, R3, 0, -R3, -R3};
    return t[e];
  }
  if (R == 4) {
    constexpr double t[4] = {0, 1, 0, -1};
    return t[e];
  }
  if (R == 3) {
    constexpr double t[3] = {0, R3, -R3};
    return t[e];
  }
  return 0.0;
}

// v *= omega_R^e (sign DIR), e a compile-time constant after unrolling
template <int R, int DIR, typename C> __host__ __device__ __forceinline__ C twc(C v, int e) {
  e = ((e % R) + R) % R;
  if (e == 0) return v;
  if (4 * e == R) return DIR > 0 ? cmuli(v) : cmulmi(v);
  if (2 * e == R) return mk<C>(-v.x, -v.y);
  if (4 * e == 3 * R) return DIR > 0 ? cmulmi(v) : cmuli(v);
  return rotc(v, cos_tab(e, R), DIR * sin_tab(e, R));
}

template <int R, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<2, DIR, C> {
  __host__ __device__ __forceinline__ static void run(C* v) {
    C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  }
};
template <int DIR, typename C> struct Dft<3, DIR, C> {
  __host__ __device__ __forceinline__ static void run(C* v) {
    using T = decltype(C::x);
    C s = cadd(v[1], v[2]);
    C d = scalec(csub(v[1], v[2]), 0.86602540378443864676);  // h d, h = sqrt(3)/2
    C a = v[0];
    v[0] = cadd(a, s);
    C m = mk<C>(a.x - (T)0.5 * s.x, a.y - (T)0.5 * s.y);
    // DIR * i * h * d
    C t = DIR > 0 ? cmuli(d) : cmulmi(d);
    v[1] = cadd(m, t);
    v[2] = csub(m, t);
  }
};
template <int DIR, typename C> struct Dft<4, DIR, C> {
  __host__ __device__ __forceinline__ static void run(C* v) {
    C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    C id13 = DIR > 0 ? cmuli(d13) : cmulmi(d13);  // DIR*i*d13
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, id13);
    v[3] = csub(d02, id13);
  }
};
// composite R = A*B:  n = A*n1 + n2, k = k1 + B*k2
template <int A, int B, int DIR, typename C> struct DftAB {
  __host__ __device__ __forceinline__ static void run(C* v) {
    constexpr int R = A * B;
    C y[A][B];
#pragma unroll
    for (int n2 = 0; n2 < A; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < B; ++n1) y[n2][n1] = v[A * n1 + n2];
      Dft<B, DIR, C>::run(y[n2]);
#pragma unroll
      for (int k1 = 0; k1 < B; ++k1) y[n2][k1] = twc<R, DIR>(y[n2][k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < B; ++k1) {
      C z[A];
#pragma unroll
      for (int n2 = 0; n2 < A; ++n2) z[n2] = y[n2][k1];
      Dft<A, DIR, C>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < A; ++k2) v[k1 + B * k2] = z[k2];
    }
  }
};
template <int DIR, typename C> struct Dft<6, DIR, C> : DftAB<2, 3, DIR, C> {};
template <int DIR, typename C> struct Dft<8, DIR, C> : DftAB<2, 4, DIR, C> {};
template <int DIR, typename C> struct Dft<12, DIR, C> : DftAB<4, 3, DIR, C> {};
template <int DIR, typename C> struct Dft<16, DIR, C> : DftAB<4, 4, DIR, C> {};

// ------------------------------------------------------- buffer accesses
// One component of the state as a raw buffer resource: 4 SGPRs of descriptor,
// a 32-bit per-lane byte offset (VGPR) and a uniform byte offset (SGPR or
// immediate).  Strided line loads then cost one VGPR for the address instead
// of a 64-bit pointer per element, and accesses past the component's size
// read zero / are dropped by the hardware.
template <typename C> struct CompBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ CompBuf(const C* base, long n) {
    unsigned long bytes = (unsigned long)n * sizeof(C);
    unsigned int nrec = bytes > 0xffffffffUL ? 0xffffffffU : (unsigned int)bytes;
    r = __builtin_amdgcn_make_buffer_rsrc(const_cast<C*>(base), (short)0, (int)nrec, 0x00020000);
  }
  __device__ __forceinline__ C ld(unsigned int voff, unsigned int soff) const {
    if constexpr (sizeof(C) == 8) {
      return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    } else {
      return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    }
  }
  __device__ __forceinline__ void st(C v, unsigned int voff, unsigned int soff) const {
    if constexpr (sizeof(C) == 8) {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, voff, soff, 0);
    } else {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, voff, soff, 0);
    }
  }
};

// ------------------------------------------------------------ line FFT plan
// radix of the stage that leaves `rem` points to combine, given Q points per
// thread: the largest of 16,12,8,6,4,3,2 dividing both.
__host__ __device__ constexpr int pick_radix(int rem, int q) {
  return (rem % 16 == 0 && q % 16 == 0) ? 16
         : (rem % 12 == 0 && q % 12 == 0) ? 12
         : (rem % 8 == 0 && q % 8 == 0)   ? 8
         : (rem % 6 == 0 && q % 6 == 0)   ? 6
         : (rem % 4 == 0 && q % 4 == 0)   ? 4
         : (rem % 3 == 0 && q % 3 == 0)   ? 3
         : (rem % 2 == 0 && q % 2 == 0)   ? 2
                                          : 0;
}
// total size of the per-stage twiddle tables (stages with NS > 1)
__host__ __device__ constexpr int twiddle_size(int n, int q) {
  int ns = 1, sz = 0;
  while (ns < n) {
    int r = pick_radix(n / ns, q);
    if (r == 0) return -1;
    if (ns > 1) sz += ns * r;
    ns *= r;
  }
  return sz;
}
// LDS index padding: one spare slot every 16 complex to spread banks
__host__ __device__ constexpr int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int line_stride(int n) { return lpad(n) + 1; }

// Barrier between the LDS writes and reads of an exchange.  WAVE: every line
// lies inside one wavefront (its P threads consecutive, P dividing 64), and
// the LDS performs one wavefront's operations in issue order, so the
// exchange only needs the compiler to keep the wave's accesses in program
// order (a wavefront-scope fence pair around a wave barrier); other waves of
// the block are not waited for.  Otherwise a block barrier.
template <bool WAVE> __device__ __forceinline__ void line_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham stage.  Thread j of the line owns butterflies b = j + u*P,
// u < Q/R; its inputs are always its registers v[u + r*Q/R] (positions
// j + P*(u + r*Q/R)), so registers hold positions j + P*r' before every stage
// and after the last one.  Intermediate outputs go through LDS.
template <int N, int Q, int NS, int DIR, typename C, bool WAVE> struct FftStage {
  static constexpr int R = pick_radix(N / NS, Q);
  static constexpr int P = N / Q;
  static constexpr bool LAST = (NS * R == N);
  static_assert(R > 0, "unsupported FFT size / points-per-thread");
  using Rw = R_t<decltype(C::x)>;
  __device__ __forceinline__ static void run(C (&v)[Q], C* lds, const Rw* tw, int j) {
    C x[Q];
    SFor<0, Q / R>::run([&](auto ui) {
      constexpr int u = decltype(ui)::value;
      const int b = j + u * P;
      const int k = b % NS;
      C y[R];
      SFor<0, R>::run([&](auto ri) {
        constexpr int r = decltype(ri)::value;
        y[r] = v[u + r * (Q / R)];
      });
      if constexpr (NS > 1) {
        SFor<1, R>::run([&](auto ri) {
          constexpr int r = decltype(ri)::value;
          y[r] = rot<DIR>(y[r], tw[r * NS + k]);
        });
      }
      Dft<R, DIR, C>::run(y);
      SFor<0, R>::run([&](auto ri) {
        constexpr int r = decltype(ri)::value;
        x[u + r * (Q / R)] = y[r];
      });
    });
    if constexpr (LAST) {
      SFor<0, Q>::run([&](auto qi) {
        constexpr int q = decltype(qi)::value;
        v[q] = x[q];
      });
    } else {
      line_sync<WAVE>();  // previous readers of this LDS line are done
      SFor<0, Q / R>::run([&](auto ui) {
        constexpr int u = decltype(ui)::value;
        const int b = j + u * P;
        const int k = b % NS;
        const int idxd = (b / NS) * NS * R + k;
        SFor<0, R>::run([&](auto ri) {
          constexpr int r = decltype(ri)::value;
          lds[lpad(idxd + r * NS)] = x[u + r * (Q / R)];
        });
      });
      line_sync<WAVE>();
      SFor<0, Q>::run([&](auto qi) {
        constexpr int q = decltype(qi)::value;
        v[q] = lds[lpad(j + P * q)];
      });
      FftStage<N, Q, NS * R, DIR, C, WAVE>::run(v, lds, tw + (NS > 1 ? NS * R : 0), j);
    }
  }
};

// In-place line FFT of N points held as v[q] = x[j + P*q] by the P threads
// of a line; afterwards v[q] = X[j + P*q].  lds: this line's LDS scratch
// (line_stride(N) elements); tw: the per-stage twiddle tables.  WAVE: see
// line_sync.
template <int N, int Q, int DIR, bool WAVE = false, typename C>
__device__ __forceinline__ void line_fft(C (&v)[Q], C* lds, const R_t<decltype(C::x)>* tw, int j) {
  FftStage<N, Q, 1, DIR, C, WAVE>::run(v, lds, tw, j);
}

// The same stages for M lines at once (v[m] the m-th line, its LDS scratch at
// lds + m * lstride): each exchange of all M lines shares one pair of
// barriers, where M calls of line_fft take M pairs.  For passes whose blocks
// hold few lines, where the barriers and LDS round trips, not the
// arithmetic, set the time.
template <int N, int Q, int NS, int DIR, int M, typename C, bool WAVE> struct FftStageM {
  static constexpr int R = pick_radix(N / NS, Q);
  static constexpr int P = N / Q;
  static constexpr bool LAST = (NS * R == N);
  static_assert(R > 0, "unsupported FFT size / points-per-thread");
  using Rw = R_t<decltype(C::x)>;
  __device__ __forceinline__ static void run(C (&v)[M][Q], C* lds, int lstride, const Rw* tw, int j) {
    C x[M][Q];
    SFor<0, M>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      SFor<0, Q / R>::run([&](auto ui) {
        constexpr int u = decltype(ui)::value;
        const int b = j + u * P;
        const int k = b % NS;
        C y[R];
        SFor<0, R>::run([&](auto ri) {
          constexpr int r = decltype(ri)::value;
          y[r] = v[m][u + r * (Q / R)];
        });
        if constexpr (NS > 1) {
          SFor<1, R>::run([&](auto ri) {
            constexpr int r = decltype(ri)::value;
            y[r] = rot<DIR>(y[r], tw[r * NS + k]);
          });
        }
        Dft<R, DIR, C>::run(y);
        SFor<0, R>::run([&](auto ri) {
          constexpr int r = decltype(ri)::value;
          x[m][u + r * (Q / R)] = y[r];
        });
      });
    });
    if constexpr (LAST) {
      SFor<0, M>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        SFor<0, Q>::run([&](auto qi) {
          constexpr int q = decltype(qi)::value;
          v[m][q] = x[m][q];
        });
      });
    } else {
      line_sync<WAVE>();  // previous readers of these LDS lines are done
      SFor<0, M>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        C* line = lds + m * lstride;
        SFor<0, Q / R>::run([&](auto ui) {
          constexpr int u = decltype(ui)::value;
          const int b = j + u * P;
          const int k = b % NS;
          const int idxd = (b / NS) * NS * R + k;
          SFor<0, R>::run([&](auto ri) {
            constexpr int r = decltype(ri)::value;
            line[lpad(idxd + r * NS)] = x[m][u + r * (Q / R)];
          });
        });
      });
      line_sync<WAVE>();
      SFor<0, M>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        const C* line = lds + m * lstride;
        SFor<0, Q>::run([&](auto qi) {
          constexpr int q = decltype(qi)::value;
          v[m][q] = line[lpad(j + P * q)];
        });
      });
      FftStageM<N, Q, NS * R, DIR, M, C, WAVE>::run(v, lds, lstride, tw + (NS > 1 ? NS * R : 0), j);
    }
  }
};
template <int N, int Q, int DIR, int M, bool WAVE = false, typename C>
__device__ __forceinline__ void line_fft_multi(C (&v)[M][Q], C* lds, int lstride, const R_t<decltype(C::x)>* tw,
                                               int j) {
  FftStageM<N, Q, 1, DIR, M, C, WAVE>::run(v, lds, lstride, tw, j);
}

// ---------------------------------------------------------------- physics
// Real-valued parameters of one run, passed by value to every kernel.
struct Phys {
  double tau0, tau1, tau2;
  double dt;
  double gx, gy, gz;
  double inv_npts;  // 1/(NX*NY*NZ), BFT normalisation (SRC/cgpe3d.f90:1782)
  int imag;         // CDT = -i DT (1) or DT (0)
  int dim;
};

__host__ __device__ __forceinline__ void sincos_t(float a, float* s, float* c) { sincosf(a, s, c); }
__host__ __device__ __forceinline__ void sincos_t(double a, double* s, double* c) { sincos(a, s, c); }
__host__ __device__ __forceinline__ float exp_t(float a) { return expf(a); }
__host__ __device__ __forceinline__ double exp_t(double a) { return exp(a); }
__host__ __device__ __forceinline__ float expm1_t(float a) { return expm1f(a); }
__host__ __device__ __forceinline__ double expm1_t(double a) { return expm1(a); }

// exp(-i CDT e) for real e: real time exp(-i DT e), imaginary time exp(-DT e)
template <typename C, typename T> __host__ __device__ __forceinline__ C phase_of(T e, T dt, int imag) {
  if (imag) return mk<C>(exp(-dt * e), (T)0);
  T s, c;
  sincos_t(dt * e, &s, &c);
  return mk<C>(c, -s);
}

// (-i CDT) * z :  real time -i DT z, imaginary time -DT z
template <typename C, typename T> __host__ __device__ __forceinline__ C mul_micdt(C z, T dt, int imag) {
  return imag ? mk<C>(-dt * z.x, -dt * z.y) : mk<C>(dt * z.y, -dt * z.x);
}

// exp(-i CDT E) for a phase theta = DT*E carried as an unevaluated sum
// hi + lo (|lo| << |hi|): the function is evaluated at hi and corrected to
// first order in lo, which is exact to rounding since lo is below an ulp of hi.
template <typename C, typename T> __host__ __device__ __forceinline__ C phase2(T hi, T lo, int imag) {
  if (imag) {
    const T e = exp_t(-hi);
    return mk<C>(e - e * lo, (T)0);
  }
  T s, c;
  sincos_t(hi, &s, &c);
  // cos(hi+lo) - i sin(hi+lo)
  return mk<C>(c - lo * s, -(s + lo * c));
}

// hi + lo = a + b exactly (Knuth's two-sum)
template <typename T> __host__ __device__ __forceinline__ void two_sum(T a, T b, T& hi, T& lo) {
  hi = a + b;
  const T bb = hi - a;
  lo = (a - (hi - bb)) + (b - bb);
}

// SP: diagonal trap + density phases (SRC/cgpe3d.f90:1023-1063), E_m formed
// from the stored amplitudes.  In FP32 the phase DT*E_m is carried as a
// two-float sum: DT*TAU0 as a two-float constant times rho with the product's
// rounding error recovered by an FMA, and exact two-sums for the trap and
// spin terms.  A stiff step (DT*TAU0*rho of order 1) then costs no more than
// the rounding of the inputs, where a plain FP32 product would perturb the
// coupling TAU0 itself by an ulp in every step.
template <int SPIN, typename C, typename T>
__host__ __device__ __forceinline__ void sp_point(C* ph, T v, const Phys& p, T dt) {
  constexpr int MS = 2 * SPIN + 1;
  T a[MS];
#pragma unroll
  for (int m = 0; m < MS; ++m) a[m] = ph[m].x * ph[m].x + ph[m].y * ph[m].y;
  T rho = 0;
#pragma unroll
  for (int m = 0; m < MS; ++m) rho += a[m];
  // per-component spin terms E_m - V - TAU0 rho (SRC/cgpe3d.f90:1030-1060)
  T es[MS];
  if constexpr (SPIN == 1) {
    const T t1 = (T)p.tau1;
    es[0] = t1 * (a[0] + a[1] - a[2]);
    es[1] = t1 * (a[0] + a[2]);
    es[2] = t1 * (a[2] + a[1] - a[0]);
  } else {
    const T t1 = (T)p.tau1, t2 = (T)p.tau2;
    const T fz = (T)2 * a[0] + a[1] - a[3] - (T)2 * a[4];
    es[0] = (T)2 * t1 * fz + (T)0.4 * t2 * a[4];
    es[1] = t1 * fz + (T)0.4 * t2 * a[3];
    es[2] = (T)0.2 * t2 * a[2];
    es[3] = -t1 * fz + (T)0.4 * t2 * a[1];
    es[4] = (T)-2 * t1 * fz + (T)0.4 * t2 * a[0];
  }
  if constexpr (sizeof(T) == 8) {
    const T base = v + (T)p.tau0 * rho;
#pragma unroll
    for (int m = 0; m < MS; ++m) ph[m] = cmul(ph[m], phase_of<C, T>(base + es[m], dt, p.imag));
  } else {
    // DT*TAU0 = c0h + c0l, DT*V and DT*es_m in FP32
    const double c0 = p.dt * p.tau0;
    const T c0h = (T)c0, c0l = (T)(c0 - (double)c0h);
    const T th = c0h * rho;
    const T tl = fma(c0h, rho, -th) + c0l * rho;
    T bh, bl;
    two_sum(th, dt * v, bh, bl);
    bl += tl;
    if (p.imag) {
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        T hi, lo;
        two_sum(bh, dt * es[m], hi, lo);
        lo += bl;
        // exp(-theta) - 1 for small theta (a factor that rounds to 1 would
        // drop the damping the same way every step), the factor itself where
        // the damping is strong
        if (hi > (T)0.25) {
          ph[m] = cmul(ph[m], phase2<C, T>(hi, lo, 1));
        } else {
          const T em1 = expm1_t(-hi);
          const T d = em1 - lo * ((T)1 + em1);
          ph[m] = cadd(ph[m], cscale(ph[m], d));
        }
      }
    } else {
      // v + v (exp(-i theta_m) - 1): the phase of a slowly changing density
      // repeats its rounding from step to step, so the factor's rounding error
      // must scale with theta, not with 1.  theta_m = theta_b + delta_m with
      // the common part theta_b = DT (V + TAU0 rho) (two-float) and the small
      // spin part delta_m = DT es_m:
      //   exp(-i theta_m) - 1 = db + dm + db dm,
      // db from one sincos of theta_b/2 (-2 sin^2 - i sin, plus the lo term),
      // dm from series when every |delta_m| <= 1/4 (two terms when <= 2^-10;
      // the truncation is below 1e-13 relative either way), else one sincos
      // each.
      T s2, c2;
      sincos_t(bh * (T)0.5, &s2, &c2);
      const T sn = (T)2 * s2 * c2, cm = (T)-2 * s2 * s2;
      const C db = mk<C>(cm - bl * sn, -(sn + bl * ((T)1 + cm)));
      T dl[MS], big = 0;
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        dl[m] = dt * es[m];
        big = fmax(big, fabs(dl[m]));
      }
      C dm[MS];
      if (big <= (T)0.0009765625) {
#pragma unroll
        for (int m = 0; m < MS; ++m) {
          const T q = dl[m] * dl[m];
          dm[m] = mk<C>(-q * (T)0.5 * ((T)1 - q * ((T)1 / 12)), -dl[m] * ((T)1 - q * ((T)1 / 6)));
        }
      } else if (big <= (T)0.25) {
#pragma unroll
        for (int m = 0; m < MS; ++m) {
          const T q = dl[m] * dl[m];
          const T sd = dl[m] * ((T)1 - q * ((T)1 / 6) * ((T)1 - q * ((T)1 / 20) * ((T)1 - q * ((T)1 / 42))));
          const T cd = -q * (T)0.5 * ((T)1 - q * ((T)1 / 12) * ((T)1 - q * ((T)1 / 30) * ((T)1 - q * ((T)1 / 56))));
          dm[m] = mk<C>(cd, -sd);
        }
      } else {
#pragma unroll
        for (int m = 0; m < MS; ++m) {
          T s3, c3;
          sincos_t(dl[m] * (T)0.5, &s3, &c3);
          dm[m] = mk<C>((T)-2 * s3 * s3, (T)-2 * s3 * c3);
        }
      }
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        const C d = cadd(cadd(db, dm[m]), cmul(db, dm[m]));
        ph[m] = cadd(ph[m], cmul(ph[m], d));
      }
    }
  }
}

// C12..C45 of SRC/cgpe3d.f90:1112-1122 at one point.  F_- = F_x - i F_y with
// the reference's SX, SY (SRC/cgpe3d.f90:686-696, standard spin-2 matrices)
// equals phi^H S_- phi, S_- = SX - i SY the lowering operator, i.e.
// 2 phi1 conj(phi2) + sqrt6 phi2 conj(phi3) + sqrt6 phi3 conj(phi4) + 2 phi4 conj(phi5)
// (1-based components; FXYZ/MAT_C at SRC/cgpe3d.f90:1086-1087,1676).
template <typename C, typename T> __host__ __device__ __forceinline__ void mat_c2(const C* ph, const Phys& p, C* c) {
  const T r = (T)1.22474487139158904910;  // sqrt(1.5)
  const C fm0 = cadd(cadd(cmulc(ph[0], ph[1]), cmulc(ph[3], ph[4])),
                     cscale(cadd(cmulc(ph[1], ph[2]), cmulc(ph[2], ph[3])), r));
  const T t1 = (T)p.tau1, t2 = (T)p.tau2;
  const C fm = cscale(fm0, (T)2 * t1);  // tau1 * F_-
  c[0] = csub(fm, cscale(cmulc(ph[3], ph[4]), (T)0.4 * t2));
  c[1] = cscale(cmulc(ph[2], ph[4]), (T)0.2 * t2);
  c[2] = csub(cscale(fm, r), cscale(cmulc(ph[2], ph[3]), (T)0.2 * t2));
  c[3] = csub(cscale(fm, r), cscale(cmulc(ph[1], ph[2]), (T)0.2 * t2));
  c[4] = cscale(cmulc(ph[0], ph[2]), (T)0.2 * t2);
  c[5] = csub(fm, cscale(cmulc(ph[0], ph[1]), (T)0.4 * t2));
}

// Hbar v with Hbar(1,2)=C12,(1,3)=C13,(2,3)=C23,(3,4)=C34,(3,5)=C35,(4,5)=C45
template <typename C> __host__ __device__ __forceinline__ void hmul2(const C* c, const C* v, C* o) {
  o[0] = cadd(cmul(c[0], v[1]), cmul(c[1], v[2]));
  o[1] = cadd(cmulc(v[0], c[0]), cmul(c[2], v[2]));
  o[2] = cadd(cadd(cmulc(v[0], c[1]), cmulc(v[1], c[2])), cadd(cmul(c[3], v[3]), cmul(c[4], v[4])));
  o[3] = cadd(cmulc(v[2], c[3]), cmul(c[5], v[4]));
  o[4] = cadd(cmulc(v[2], c[4]), cmulc(v[3], c[5]));
}

// acc + a b and acc + a conj(b)
template <typename C> __host__ __device__ __forceinline__ C cmac(C acc, C a, C b) {
  return mk<C>(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
template <typename C> __host__ __device__ __forceinline__ C cmacc(C acc, C a, C b) {
  return mk<C>(fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(a.y, b.x, fma(-a.x, b.y, acc.y)));
}

// hmul2 with every product accumulated by FMA (12 complex products, 48 FMAs)
template <typename C> __host__ __device__ __forceinline__ void hmul2f(const C* c, const C* v, C* o) {
  o[0] = cmac(cmul(c[0], v[1]), c[1], v[2]);
  o[1] = cmac(cmulc(v[0], c[0]), c[2], v[2]);
  o[2] = cmac(cmac(cmacc(cmulc(v[0], c[1]), v[1], c[2]), c[3], v[3]), c[4], v[4]);
  o[3] = cmac(cmulc(v[2], c[3]), c[5], v[4]);
  o[4] = cmacc(cmulc(v[2], c[4]), v[3], c[5]);
}

// 1/k for a small positive integer k (one correctly rounded or 1-ulp value;
// the terms it scales are far below the rounding of the state)
__host__ __device__ __forceinline__ float recip_k(int k) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf((float)k);
#else
  return 1.0f / (float)k;
#endif
}

// SE spin-2 (SRC/cgpe3d.f90:1673-1741): predictor with factor cpred (1/2 in
// 1D, 1 in 2D/3D), averaged couplings, then exp(-i CDT Hbar) phi.  ZHEEV's
// V diag(exp(-i CDT w)) V^H is replaced by its Taylor series applied to the
// vector, summed to rounding (the matrix exponential is unique) apart from the
// zeroth term, which is added last so rounding scales with the change; when
// |CDT|*||Hbar||_inf exceeds 1/2 the exponential is split into equal
// sub-steps.  FP64 stops the series when a term falls below tol relative to
// the vector.  FP32 fixes the number of terms K beforehand from the bound
// bs = |CDT| h ||Hbar||_inf (h the sub-step): the remainder after K terms is
// at most bs^(K+1)/(K+1)! e^bs ||phi||, and K is the smallest with
// bs^(K+1)/(K+1)! <= 2^-34, far below FP32 rounding, and a term costs one
// sparse matrix-vector product with fused multiply-adds and one scaling.
template <typename C, typename T>
__host__ __device__ __forceinline__ void se2_point(C* ph, const Phys& p, T dt, T cpred, T tol) {
  constexpr bool F32 = sizeof(T) == 4;
  C c0[6], c1[6], hv[5], p1[5];
  mat_c2<C, T>(ph, p, c0);
  if constexpr (F32)
    hmul2f(c0, ph, hv);
  else
    hmul2(c0, ph, hv);
#pragma unroll
  for (int m = 0; m < 5; ++m) p1[m] = cadd(ph[m], mul_micdt<C, T>(hv[m], dt * cpred, p.imag));
  mat_c2<C, T>(p1, p, c1);
  // averaged couplings; FP32 keeps the sum and folds the 1/2 into the time
  // factor (an exact power of two either way)
  const T half = F32 ? (T)0.5 : (T)1;
#pragma unroll
  for (int i = 0; i < 6; ++i) c0[i] = F32 ? cadd(c0[i], c1[i]) : cscale(cadd(c0[i], c1[i]), (T)0.5);
  // infinity-norm bound with |re|+|im|
  T n[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) n[i] = fabs(c0[i].x) + fabs(c0[i].y);
  T rows = fmax(fmax(n[0] + n[1], n[0] + n[2]), fmax(n[1] + n[2] + n[3] + n[4], fmax(n[3] + n[5], n[4] + n[5])));
  T b = dt * half * rows;
  // bounded sub-step count (a non-finite bound must not spin the loop)
  int nsub = b > (T)0.5 ? (b < (T)32 ? (int)ceil(b / (T)0.5) : 64) : 1;
  T h = dt, bs = b;  // per sub-step
  if (nsub > 1) {
    h = dt / (T)nsub;
    bs = b / (T)nsub;
  }
  if constexpr (sizeof(T) == 4) {
    // K terms suffice when bs <= t_K = ((K+1)! 2^-34)^(1/(K+1)) (rounded
    // down); past t_9 the bound is followed term by term
    constexpr float tk[9] = {1.077880e-05f, 7.035192e-04f, 6.107498e-03f, 2.335596e-02f, 5.887771e-02f,
                             1.165034e-01f, 1.976415e-01f, 3.020595e-01f, 4.285878e-01f};
    int K = 1;
#pragma unroll
    for (int i = 0; i < 9; ++i) K += bs > tk[i] ? 1 : 0;
    if (bs > tk[8]) {
      K = 1;
      T rem = bs * bs * (T)0.5;
      while (rem > (T)5.8207660913467407e-11 && K < 30) {
        ++K;
        rem *= bs * recip_k(K + 1);
      }
    }
    const T hh = h * half;
    for (int s = 0; s < nsub; ++s) {
      C term[5], sum[5];
      hmul2f(c0, ph, hv);
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        term[m] = mul_micdt<C, T>(hv[m], hh, p.imag);
        sum[m] = term[m];
      }
      for (int k = 2; k <= K; ++k) {
        hmul2f(c0, term, hv);
        const T f = hh * recip_k(k);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          term[m] = mul_micdt<C, T>(hv[m], f, p.imag);
          sum[m] = cadd(sum[m], term[m]);
        }
      }
#pragma unroll
      for (int m = 0; m < 5; ++m) ph[m] = cadd(ph[m], sum[m]);
    }
    return;
  }
  for (int s = 0; s < nsub; ++s) {
    C term[5], sum[5];
    T smax = 0;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      term[m] = ph[m];
      sum[m] = mk<C>((T)0, (T)0);
      smax = fmax(smax, fabs(ph[m].x) + fabs(ph[m].y));
    }
    for (int k = 1; k < 40; ++k) {
      hmul2(c0, term, hv);
      T f = h / (T)k;
      T tmax = 0;
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        term[m] = mul_micdt<C, T>(hv[m], f, p.imag);
        sum[m] = cadd(sum[m], term[m]);
        tmax = fmax(tmax, fabs(term[m].x) + fabs(term[m].y));
      }
      if (tmax <= tol * smax) break;
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) ph[m] = cadd(ph[m], sum[m]);
  }
}

// complex cos and sin of w = a + i b
template <typename C, typename T> __host__ __device__ __forceinline__ void ccossin(C w, C& cw, C& sw) {
  T sa, ca;
  sincos_t(w.x, &sa, &ca);
  T ch = cosh(w.y), sh = sinh(w.y);
  cw = mk<C>(ca * ch, -sa * sh);
  sw = mk<C>(sa * ch, ca * sh);
}
template <typename C> __host__ __device__ __forceinline__ C cdiv(C a, C b) {
  auto d = b.x * b.x + b.y * b.y;
  return mk<C>((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// SE spin-1, closed form (SRC/cgpe3d.f90:1637-1672).  For |omega| below
// `small` the functions (cos w - 1)/w^2 and -i sin(w)/w come from their
// Taylor series (the closed form loses all digits there in FP32); exactly at
// omega = 0 both forms give delta = -1/2, kappa = -i as the reference does.
template <typename C, typename T>
__host__ __device__ __forceinline__ void se1_point(C* ph, const Phys& p, T dt, T small) {
  const T t1 = (T)p.tau1;
  C a1 = cscale(cmulc(ph[1], ph[2]), t1);
  C b1 = cscale(cmulc(ph[1], ph[0]), t1);
  T om_r = sqrt(cabs2(a1) + cabs2(b1));
  // CDT: real time (dt,0), imaginary (0,-dt)
  C cdt = p.imag ? mk<C>((T)0, -dt) : mk<C>(dt, (T)0);
  C w = cscale(cdt, om_r);
  C delta, kappa;
  T aw = sqrt(cabs2(w));
  if (aw > small) {
    C cw, sw;
    ccossin<C, T>(w, cw, sw);
    C w2 = cmul(w, w);
    delta = cdiv(mk<C>(cw.x - (T)1, cw.y), w2);
    kappa = cmulmi(cdiv(sw, w));
  } else {
    C w2 = cmul(w, w);
    C w4 = cmul(w2, w2);
    C w6 = cmul(w4, w2);
    delta = mk<C>((T)-0.5 + w2.x / (T)24 - w4.x / (T)720 + w6.x / (T)40320,
                  w2.y / (T)24 - w4.y / (T)720 + w6.y / (T)40320);
    C sk = mk<C>((T)1 - w2.x / (T)6 + w4.x / (T)120 - w6.x / (T)5040,
                 -w2.y / (T)6 + w4.y / (T)120 - w6.y / (T)5040);
    kappa = cmulmi(sk);
  }
  C dc2 = cmul(delta, cmul(cdt, cdt));
  C kc = cmul(kappa, cdt);
  C aa = mk<C>(cabs2(a1), (T)0), bb = mk<C>(cabs2(b1), (T)0);
  C t1v = cadd(cmul(dc2, cadd(cmul(aa, ph[0]), cmul(cmulc(a1, b1), ph[2]))), cmul(kc, cmul(a1, ph[1])));
  C t2v = cadd(cmul(dc2, cmul(cadd(aa, bb), ph[1])), cmul(kc, cadd(cmulc(ph[0], a1), cmulc(ph[2], b1))));
  C t3v = cadd(cmul(dc2, cadd(cmul(cmulc(b1, a1), ph[0]), cmul(bb, ph[2]))), cmul(kc, cmul(b1, ph[1])));
  ph[0] = cadd(ph[0], t1v);
  ph[1] = cadd(ph[1], t2v);
  ph[2] = cadd(ph[2], t3v);
}

// SOC factor A(beta) diag(exp(-i CDT L alpha)) A(beta)^H at one k-point
// (SRC/cgpe3d.f90:904-1004).  A(beta) = E A0 with A0 the real constant matrix
// of the 1D code (SRC/cgpe1d.f90:627-675) and E = diag(e1^S, ..., e1^-S),
// e1 = exp(-i beta).  Returned is only the change of v,
//   corr = E A0 diag(dm1) A0^T E^H v,   dm1[c] = exp(-i CDT L alpha) - 1,
// so that v + corr is the factor applied to v: rounding then scales with the
// (small, for the low modes that carry the state) change rather than with v,
// and at alpha = 0 the result is v exactly.
template <int SPIN, typename C, typename T>
__host__ __device__ __forceinline__ void soc_corr(const C* v, C e1, const C* dm1, C* corr) {
  if constexpr (SPIN == 1) {
    const T s5 = (T)0.70710678118654752440;
    // A0 rows: [.5,-s5,.5],[-s5,0,s5],[.5,s5,.5]
    C e2 = cmul(e1, e1);
    C w0 = cmulc(v[0], e2), w1 = cmulc(v[1], e1), w2 = v[2];  // E^H v
    // A0^T w
    C u0 = cadd(cscale(cadd(w0, w2), (T)0.5), cscale(w1, -s5));
    C u1 = cscale(csub(w2, w0), s5);
    C u2 = cadd(cscale(cadd(w0, w2), (T)0.5), cscale(w1, s5));
    u0 = cmul(u0, dm1[0]);
    u1 = cmul(u1, dm1[1]);
    u2 = cmul(u2, dm1[2]);
    // A0 u
    C r0 = csub(cscale(cadd(u0, u2), (T)0.5), cscale(u1, s5));
    C r1 = cscale(csub(u2, u0), s5);
    C r2 = cadd(cscale(cadd(u0, u2), (T)0.5), cscale(u1, s5));
    corr[0] = cmul(r0, e2);
    corr[1] = cmul(r1, e1);
    corr[2] = r2;
  } else {
    const T q = (T)0.61237243569579452455;  // sqrt(0.375)
    C e2 = cmul(e1, e1);
    C w0 = cmulc(v[0], e2), w1 = cmulc(v[1], e1), w2 = v[2], w3 = cmul(v[3], e1), w4 = cmul(v[4], e2);
    // A0 rows (1D constant A): r1 [.25,-.5,q,-.5,.25], r2 [-.5,.5,0,-.5,.5],
    // r3 [q,0,-.5,0,q], r4 [-.5,-.5,0,.5,.5], r5 [.25,.5,q,.5,.25]
    // u = A0^T w (u_c = sum_r A0[r][c] w_r)
    C s04 = cadd(w0, w4), s13 = cadd(w1, w3), d13 = csub(w3, w1), d04 = csub(w4, w0);
    C u0 = cadd(cadd(cscale(s04, (T)0.25), cscale(s13, (T)-0.5)), cscale(w2, q));
    C u1 = cadd(cscale(d04, (T)0.5), cscale(csub(w1, w3), (T)0.5));
    C u2 = cadd(cscale(s04, q), cscale(w2, (T)-0.5));
    C u3 = cadd(cscale(d04, (T)0.5), cscale(d13, (T)0.5));
    C u4 = cadd(cadd(cscale(s04, (T)0.25), cscale(s13, (T)0.5)), cscale(w2, q));
    u0 = cmul(u0, dm1[0]);
    u1 = cmul(u1, dm1[1]);
    u2 = cmul(u2, dm1[2]);
    u3 = cmul(u3, dm1[3]);
    u4 = cmul(u4, dm1[4]);
    // r = A0 u
    C r0 = cadd(cadd(cscale(u0, (T)0.25), cscale(cadd(u1, u3), (T)-0.5)), cadd(cscale(u2, q), cscale(u4, (T)0.25)));
    C r1 = cscale(cadd(csub(u1, u0), csub(u4, u3)), (T)0.5);
    C r2 = cadd(cscale(cadd(u0, u4), q), cscale(u2, (T)-0.5));
    C r3 = cscale(cadd(csub(u3, u0), csub(u4, u1)), (T)0.5);
    C r4 = cadd(cadd(cscale(u0, (T)0.25), cscale(cadd(u1, u3), (T)0.5)), cadd(cscale(u2, q), cscale(u4, (T)0.25)));
    corr[0] = cmul(r0, e2);
    corr[1] = cmul(r1, e1);
    corr[2] = r2;
    corr[3] = cmulc(r3, e1);
    corr[4] = cmulc(r4, e2);
  }
}


// e1 = exp(-i beta) and d[L+S] = exp(-i CDT L alpha) for one (kx, ky)
// (ALPHA, BETA as in SRC/cgpe3d.f90:956-961; beta = 0 when alpha <= 1e-40).
// In 1D the reference uses the constant basis with exponent 2 L GAMMAX kx
// (SRC/cgpe1d.f90:1518), i.e. e1 = 1 and a signed alpha = 2 GAMMAX kx.
template <int SPIN, typename C, typename T>
__host__ __device__ __forceinline__ void soc_factors(T kx, T ky, const Phys& p, T dt, C& e1, C* dm1) {
  T alpha;
  if (p.dim == 1) {
    alpha = (T)2 * (T)p.gx * kx;
    e1 = mk<C>((T)1, (T)0);
  } else {
    T ax = (T)p.gx * kx, ay = (T)p.gy * ky;
    T r = sqrt(ax * ax + ay * ay);
    alpha = (T)2 * r;
    if (alpha <= (T)1e-30)
      e1 = mk<C>((T)1, (T)0);
    else
      e1 = mk<C>(ax / r, -ay / r);
  }
  dm1[SPIN] = mk<C>((T)0, (T)0);
#pragma unroll
  for (int L = 1; L <= SPIN; ++L) {
    const T th = dt * (T)L * alpha;
    if (p.imag) {
      dm1[SPIN + L] = mk<C>(expm1_t(-th), (T)0);
      dm1[SPIN - L] = mk<C>(expm1_t(th), (T)0);
    } else {
      // exp(-i th) - 1 = -2 sin^2(th/2) - i sin(th)
      T sh, ch;
      sincos_t(th * (T)0.5, &sh, &ch);
      const T re = (T)-2 * sh * sh, im = (T)2 * sh * ch;
      dm1[SPIN + L] = mk<C>(re, -im);
      dm1[SPIN - L] = mk<C>(re, im);
    }
  }
}

// e1 and dm1 of soc_factors from one entry of the 2D table (built on the host
// in FP64 by soc2_table): t[0] = e1, t[L] packs the two factors of +-L --
// real time (re, im) of exp(-i th) - 1 = (re, -im) for +L and its conjugate
// for -L; imaginary time (expm1(-th), expm1(th)).  Forming the factors per
// point in the 2D k pass (a square root, a division and a sincos per point)
// held 250-290 VGPRs; the table read holds about 110.
template <int SPIN, typename C>
__host__ __device__ __forceinline__ void soc_from_table(const C* t, int imag, C& e1, C* dm1) {
  using T = decltype(C::x);
  e1 = t[0];
  dm1[SPIN] = mk<C>((T)0, (T)0);
#pragma unroll
  for (int L = 1; L <= SPIN; ++L) {
    const C d = t[L];
    if (imag) {
      dm1[SPIN + L] = mk<C>(d.x, (T)0);
      dm1[SPIN - L] = mk<C>(d.y, (T)0);
    } else {
      dm1[SPIN + L] = mk<C>(d.x, -d.y);
      dm1[SPIN - L] = mk<C>(d.x, d.y);
    }
  }
}

// soc_from_table for wave-number indices (x, y) of an nx x ny grid, from the
// quarter table of soc2_table: the entry of (|kx|, |ky|) (FFT index i and
// n - i hold opposite wave numbers; i <= n/2 is the non-negative one), with
// the real part of e1 negated for kx < 0 and the imaginary part for ky < 0.
template <int SPIN, typename C>
__host__ __device__ __forceinline__ void soc_from_quarter(const C* t, int x, int y, int nx, int ny, int imag, C& e1,
                                                          C* dm1) {
  const bool nxneg = x > nx / 2, nyneg = y > ny / 2;
  const int xq = nxneg ? nx - x : x, yq = nyneg ? ny - y : y;
  soc_from_table<SPIN, C>(t + ((long)yq * (nx / 2 + 1) + xq) * (SPIN + 1), imag, e1, dm1);
  if (nxneg) e1.x = -e1.x;
  if (nyneg) e1.y = -e1.y;
}

}  // namespace fgpe
