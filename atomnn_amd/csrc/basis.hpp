// Device forms of the reference's basis functions and activations beyond GAUSSIAN / GAUSS /
// tanh: the value and the derivative each kernel needs, in FP64 or FP32.
//
//   radial  phi(x), x = -eta (r - rs)            basis_radial.cpp:302-359 (values), 371-438
//   angular distance factor h(x), x = ietap r^alpha, per leg (separable) or per pair
//                                                basis_angular.cpp:415-441, 478-630
//   activations a(z), d(z)                       ml/nn.cpp:144-766 (AFFPBP)
//   log(1 + x) as the reference computes it       math/special.cpp:131-135
//   sech                                         math/special.cpp:71-79
#pragma once

#include "device.hpp"
#include "model.hpp"

namespace atomnn {

// The reference's log(1 + x): 2 atanh(y), y = x / (x + 2), odd series through y^9.
template <typename T>
__device__ __forceinline__ T logp1_ref(T x) {
    const T y = x / (x + T(2));
    const T y2 = y * y;
    return T(2) * y * (T(1) + y2 * (T(1.0 / 3.0) + y2 * (T(1.0 / 5.0) + y2 * (T(1.0 / 7.0) + y2 * T(1.0 / 9.0)))));
}

template <typename T>
__device__ __forceinline__ T sech_ref(T x) {
    const T e = exp(-fabs(x));
    return T(2) * e / (T(1) + e * e);
}

// GAUSS, IPOWP, SECHP factorise over the two legs of a triplet; IPOWS and SECHS do not
__device__ __forceinline__ bool ang_separable(int kind) { return kind <= ANG_SECHP; }

// the reference's ln 2 constant (math/const.hpp:15)
constexpr double kLog2Ref = 0.693147180559945;

// Radial function value v and dv = dphi/dx for the non-Gaussian kinds.  SWISH and MISH at
// x = 0 exactly take their limits (the reference leaves the previous value there).
template <typename T>
__device__ __forceinline__ void rad_phi(int kind, T x, T& v, T& dv) {
    switch (kind) {
        case RAD_TANH: {
            const T t = tanh(x);
            v = T(0.5) * (t + T(1));
            dv = T(0.5) * (T(1) - t * t);
        } break;
        case RAD_SOFTPLUS: {
            const T e = exp(-fabs(x));
            const T inv = T(1) / (T(1) + e);
            if (x > T(0)) {
                v = x + logp1_ref(e);
                dv = inv;
            } else {
                v = logp1_ref(e);
                dv = e * inv;
            }
        } break;
        case RAD_LOGCOSH: {
            const T e = exp(T(-2) * fabs(x));
            const T inv = T(1) / (T(1) + e);
            if (x > T(0)) {
                v = x + T(0.5) * logp1_ref(e);
                dv = inv;
            } else {
                v = T(0.5) * logp1_ref(e);
                dv = e * inv;
            }
        } break;
        case RAD_SWISH: {
            const T e = exp(-fabs(x));
            const T den = T(1) / (T(1) + e);
            if (x > T(0)) {
                v = x * den;
                dv = (T(1) + e * (T(1) + x)) * den * den;
            } else if (x < T(0)) {
                v = x * e * den;
                dv = (T(1) + x + e) * e * den * den;
            } else {
                v = T(0);
                dv = T(0.5);
            }
        } break;
        default: {  // RAD_MISH
            const T e = exp(-fabs(x));
            if (x > T(0)) {
                const T den = T(1) / (T(2) * e * (e + T(1)) + T(1));
                v = x * (T(2) * e + T(1)) * den;
                dv = (T(1) + e * (T(4) + e * (T(4) * x + T(6) + e * T(4) * (x + T(1))))) * den * den;
            } else if (x < T(0)) {
                const T den = T(1) / (e * (e + T(2)) + T(2));
                v = x * e * (e + T(2)) * den;
                dv = (e * (e * (e + T(4)) + T(4) * x + T(6)) + T(4) * (x + T(1))) * den * den * e;
            } else {
                v = T(0);
                dv = T(0.6);
            }
        } break;
    }
}

// Separable angular leg factor h(x) and dh/dx: GAUSS e^-x, IPOWP 1/(1+x), SECHP sech x
// (tanh taken as sqrt(1 - sech^2), as the reference does).
template <typename T>
__device__ __forceinline__ void ang_leg(int kind, T x, T& h, T& dh) {
    if (kind == ANG_GAUSS) {
        h = exp(-x);
        dh = -h;
    } else if (kind == ANG_IPOWP) {
        h = T(1) / (T(1) + x);
        dh = -h * h;
    } else {  // ANG_SECHP
        h = sech_ref(x);
        dh = -h * sqrt(T(1) - h * h);
    }
}

// Per-entry quantity a(x) of a non-separable block (IPOWS: x; SECHS: e^-x) and the pair factor
// F of two legs with t such that dF/dx_j = -t F:  IPOWS F = 1/(1 + x_j + x_k), t = F;
// SECHS F = sech(x_j + x_k) = 2 a_j a_k / (1 + (a_j a_k)^2), t = sqrt(1 - F^2).
template <typename T>
__device__ __forceinline__ T ang_entry(int kind, T x) {
    return kind == ANG_IPOWS ? x : exp(-x);
}
template <typename T>
__device__ __forceinline__ void ang_pair(int kind, T aj, T ak, T& F, T& t) {
    if (kind == ANG_IPOWS) {
        F = T(1) / (T(1) + aj + ak);
        t = F;
    } else {
        const T ee = aj * ak;
        F = T(2) * ee / (T(1) + ee * ee);
        t = sqrt(T(1) - F * F);
    }
}

// tanh z = -m / (2 + m) with m = expm1(-2|z|), sign of z: within 3 ulp of the library tanh
// (expm1 keeps the relative accuracy near 0) in about a third of its FP64 instructions.  The
// tanh-only network, fused and training kernels use it; FP32 keeps the library tanhf.
__device__ __forceinline__ double tanh_m1(double z) {
    const double m = expm1(-2.0 * fabs(z));
    return copysign(-m / (2.0 + m), z);
}
__device__ __forceinline__ float tanh_m1(float z) { return tanhf(z); }

// Hidden-layer activation a(z) and derivative d(z) (AFFPBP, ml/nn.cpp:144-766).
template <typename T>
__device__ __forceinline__ void act_eval(int kind, T z, T& a, T& d) {
    switch (kind) {
        case NEU_TANH: {
            a = tanh(z);
            d = T(1) - a * a;
        } return;
        case NEU_LINEAR: {
            a = z;
            d = T(1);
        } return;
        case NEU_SIGMOID: {
            const T e = exp(-fabs(z)), fr = T(1) / (T(1) + e);
            a = z >= T(0) ? fr : e * fr;
            d = e * fr * fr;
        } return;
        case NEU_ISRU: {
            const T r = T(1) / sqrt(T(1) + z * z);
            a = z * r;
            d = r * r * r;
        } return;
        case NEU_ARCTAN: {
            const T c = T(2.0 / 3.141592653589793);
            a = c * atan(z);
            d = c / (T(1) + z * z);
        } return;
        case NEU_SOFTSIGN: {
            const T r = T(1) / (T(1) + fabs(z));
            a = z * r;
            d = r * r;
        } return;
        case NEU_RELU: {
            a = z > T(0) ? z : T(0);
            d = z > T(0) ? T(1) : T(0);
        } return;
        case NEU_SOFTPLUS: {
            const T e = exp(-fabs(z)), inv = T(1) / (T(1) + e);
            if (z >= T(0)) {
                a = z + logp1_ref(e) - T(kLog2Ref);
                d = inv;
            } else {
                a = logp1_ref(e) - T(kLog2Ref);
                d = e * inv;
            }
        } return;
        case NEU_ELU: {
            if (z >= T(0)) {
                a = z;
                d = T(1);
            } else {
                const T e = exp(z);
                a = e - T(1);
                d = e;
            }
        } return;
        case NEU_GELU:
        case NEU_SWISH: {
            const T X = kind == NEU_GELU ? T(1.702) : T(1);
            const T e = exp(-fabs(X * z)), den = T(1) / (T(1) + e);
            if (z > T(0)) {
                a = z * den;
                d = (T(1) + e * (T(1) + X * z)) * den * den;
            } else {
                a = z * e * den;
                d = e * (X * z + T(1) + e) * den * den;
            }
        } return;
        case NEU_MISH: {
            const T e = exp(-fabs(z));
            if (z > T(0)) {
                const T den = T(1) / (T(2) * e * (e + T(1)) + T(1));
                a = z * (T(2) * e + T(1)) * den;
                d = (T(1) + e * (T(4) + e * (T(4) * z + T(6) + e * T(4) * (z + T(1))))) * den * den;
            } else {
                const T den = T(1) / (e * (e + T(2)) + T(2));
                a = z * e * (e + T(2)) * den;
                d = e * (e * (e * (e + T(4)) + T(4) * z + T(6)) + T(4) * (z + T(1))) * den * den;
            }
        } return;
        case NEU_TANHRE: {
            if (z > T(0)) {
                a = z;
                d = T(1);
            } else {
                a = tanh(z);
                d = T(1) - a * a;
            }
        } return;
        case NEU_LOGCOSH: {
            const T e = exp(T(-2) * fabs(z)), inv = T(1) / (e + T(1));
            if (z >= T(0)) {
                a = z + T(0.5) * (logp1_ref(e) - T(kLog2Ref));
                d = inv;
            } else {
                a = T(0.5) * (logp1_ref(e) - T(kLog2Ref));
                d = e * inv;
            }
        } return;
        case NEU_IQUAD: {
            const T r = T(1) / sqrt(T(1) + z * z);
            a = T(0.5) * z * (T(1) + z * r);
            d = T(0.5) * (T(1) + z * (z * z + T(2)) * r * r * r);
        } return;
        case NEU_IFABS: {
            if (z > T(0)) {
                const T r = T(1) / (T(1) + z);
                a = T(0.5) * z * (T(1) + z * r);
                d = T(1) - T(0.5) * r * r;
            } else {
                const T r = T(1) / (T(1) - z);
                a = T(0.5) * z * (T(1) + z * r);
                d = T(0.5) * r * r;
            }
        } return;
        default: {  // NEU_SINSIG
            const T pi = T(3.141592653589793);
            const T e = exp(-z), den = T(1) / (T(1) + e);
            const T sv = sin(T(0.5) * pi * den), cv = cos(T(0.5) * pi * den);
            a = z * sv;
            d = sv + T(0.5) * pi * z * e * cv * den * den;
        } return;
    }
}

// Second derivative a''(z) of the hidden-layer activations (AFFPBP2, ml/nn.cpp:149-766), the
// reference's expressions where they are the derivative of its d(z); for SOFTSIGN (left unset
// there), IQUAD, IFABS and SINSIG (set to 0 there) the true second derivative of its a(z).
__device__ __forceinline__ double act_d2(int kind, double z) {
    switch (kind) {
        case NEU_TANH:
        case NEU_TANHRE: {
            if (kind == NEU_TANHRE && z > 0.0) return 0.0;
            const double t = tanh(z);
            return -2.0 * t * (1.0 - t * t);
        }
        case NEU_SIGMOID: {
            const double e = exp(-fabs(z)), fr = 1.0 / (1.0 + e);
            return (z >= 0.0 ? -e * (1.0 - e) : -e * (e - 1.0)) * fr * fr * fr;
        }
        case NEU_ISRU: {
            const double r = 1.0 / sqrt(1.0 + z * z), r2 = r * r;
            return -3.0 * z * r2 * r2 * r;
        }
        case NEU_ARCTAN: {
            const double fr = 1.0 / (1.0 + z * z);
            return -2.0 * z * (2.0 / 3.141592653589793) * fr * fr;
        }
        case NEU_SOFTSIGN: {
            const double r = 1.0 / (1.0 + fabs(z));
            return (z > 0.0 ? -2.0 : (z < 0.0 ? 2.0 : 0.0)) * r * r * r;
        }
        case NEU_SOFTPLUS: {
            const double e = exp(-fabs(z)), fr = 1.0 / (1.0 + e);
            return e * fr * fr;
        }
        case NEU_ELU: return z >= 0.0 ? 0.0 : exp(z);
        case NEU_GELU:
        case NEU_SWISH: {
            const double X = kind == NEU_GELU ? 1.702 : 1.0;
            const double e = exp(-fabs(X * z)), den = 1.0 / (1.0 + e), xz = X * z;
            return z > 0.0 ? X * e * (2.0 - xz + e * (2.0 + xz)) * den * den * den
                           : X * e * (2.0 + xz + e * (2.0 - xz)) * den * den * den;
        }
        case NEU_MISH: {
            const double e = exp(-fabs(z));
            if (z > 0.0) {
                const double den = 1.0 / (2.0 * e * (e + 1.0) + 1.0);
                return 4.0 * e * e * (e * (2.0 * e * (e * (z + 2.0) + z + 4.0) + 3.0 * (-z + 2.0)) - 2.0 * (z - 1.0)) * den *
                       den * den;
            }
            const double den = 1.0 / (e * (e + 2.0) + 2.0);
            return 4.0 * e * (2.0 * (z + 2.0) + e * (e * (-2.0 * e * (z - 1.0) - 3.0 * (z - 2.0)) + 2.0 * (z + 4.0))) * den *
                   den * den;
        }
        case NEU_LOGCOSH: {
            const double e = exp(-2.0 * fabs(z)), den = 1.0 / (e + 1.0);
            return 2.0 * e * den * den;
        }
        case NEU_IQUAD: {
            const double r = 1.0 / sqrt(1.0 + z * z), r2 = r * r;
            return 0.5 * (2.0 - z * z) * r2 * r2 * r;
        }
        case NEU_IFABS: {
            const double r = 1.0 / (1.0 + fabs(z));
            return r * r * r;
        }
        case NEU_SINSIG: {
            const double pi = 3.141592653589793;
            const double e = exp(-z), s = 1.0 / (1.0 + e), s1 = e * s * s, s2 = s1 * (1.0 - 2.0 * s);
            const double sv = sin(0.5 * pi * s), cv = cos(0.5 * pi * s);
            return pi * s1 * cv + z * 0.5 * pi * (cv * s2 - 0.5 * pi * s1 * s1 * sv);
        }
        default: return 0.0;  // NEU_LINEAR, NEU_RELU
    }
}

}  // namespace atomnn
