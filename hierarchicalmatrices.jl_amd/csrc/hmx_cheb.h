// hmx_cheb.h -- chebyshevpoints(T, n) and chebyshevbarycentricweights(T, n)
// (src/BarycentricMatrix.jl:92-136) in the element type T:
// the argument of sinpi is formed in T, sin(pi q) is evaluated in long double
// and rounded once to T.
#pragma once
#include <algorithm>
#include <cmath>

template <class T>
inline void cheb_points_t(T *x, int n, int kind = 1) {
    const long double PI_L = 3.14159265358979323846264338327950288419716939937510L;
    std::fill(x, x + n, T(0));
    const T den = kind == 1 ? (T)(2 * n) : (T)(2 * (n - 1));
    for (int k = 1; k <= n / 2; ++k) {
        const T q = ((T)(n - 2 * k) + T(1)) / den;
        x[k - 1] = (T)sinl(PI_L * (long double)q);
    }
    for (int k = 1; k <= n / 2; ++k) x[n - k] = -x[k - 1];
}

template <class T>
inline void cheb_weights_t(T *l, int n, int kind = 1) {
    const long double PI_L = 3.14159265358979323846264338327950288419716939937510L;
    std::fill(l, l + n, T(0));
    if (kind == 1) {
        for (int k = 1; k <= n / 2 + 1 && k <= n; ++k) {
            const T q = ((T)(2 * k) - T(1)) / (T)(2 * n);
            l[k - 1] = (T)sinl(PI_L * (long double)q);
        }
        for (int k = 1; k <= n / 2; ++k) l[n - k] = l[k - 1];
    } else {
        std::fill(l, l + n, T(1));
    }
    for (int k = 2; k <= n; k += 2) l[k - 1] = -l[k - 1];
    if (kind != 1) {
        l[0] *= T(0.5);
        l[n - 1] *= T(0.5);
    }
}
