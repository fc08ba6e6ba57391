// Device helpers of the Gauss-point coefficient evaluation, shared by the
// interpreter (assemble.hip) and the per-program kernels compiled at plan
// creation (coefjit.cpp embeds this header in their source).
#pragma once
#include "device.h"

namespace slx {

// Jacobian, inverse and determinant of one (element, Gauss point).  Every
// index below is a compile-time constant (loops fully unrolled, dimensions
// guarded), so the values stay in registers: the program reads them through
// a constant-index switch instead of indexing a private array.
struct Geo {
    double J[9], inv[9], det;
};

__device__ inline double pick9(const double* v, int i) {
    switch (i) {
        case 0: return v[0];
        case 1: return v[1];
        case 2: return v[2];
        case 3: return v[3];
        case 4: return v[4];
        case 5: return v[5];
        case 6: return v[6];
        case 7: return v[7];
        case 8: return v[8];
    }
    return 0.0;
}

// Corner-node coordinates of one element, loaded once and reused for every
// Gauss point the thread evaluates.
// Curved elements (more than SLX_MAX_GEO geometry nodes) keep only the
// element's node list and read the coordinates per Gauss point.
struct GeoX {
    double x[SLX_MAX_GEO][3];
    const int* nodes;
    const double* xyz;
};

__device__ inline void geo_load(const DevBucket& b, const double* xyz, int e, GeoX& gx) {
    gx.nodes = b.geo_nodes + (size_t)e * b.n_geo;
    gx.xyz = xyz;
    if (b.n_geo > SLX_MAX_GEO) {
#pragma unroll
        for (int k = 0; k < SLX_MAX_GEO; k++) gx.x[k][0] = gx.x[k][1] = gx.x[k][2] = 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < SLX_MAX_GEO; k++) {
        if (k < b.n_geo) {
            const int node = b.geo_nodes[(size_t)e * b.n_geo + k];
            gx.x[k][0] = xyz[3 * (size_t)node];
            gx.x[k][1] = xyz[3 * (size_t)node + 1];
            gx.x[k][2] = xyz[3 * (size_t)node + 2];
        } else {
            gx.x[k][0] = gx.x[k][1] = gx.x[k][2] = 0.0;
        }
    }
}

// J[r][c] = sum_k x_k^c dN_k/dki_r (rawfield.cpp:2442-2502); det and inverse
// as jacobian.cpp:308-349, 396-505.
__device__ inline void geo_eval(const DevBucket& b, const GeoX& gx, int q, Geo& g) {
    const auto& x = gx.x;
    if (b.n_geo > SLX_MAX_GEO) {
        // curved: J[r][c] summed node by node (same order over k as below)
#pragma unroll
        for (int i = 0; i < 9; i++) g.J[i] = 0.0;
        for (int k = 0; k < b.n_geo; k++) {
            const int node = gx.nodes[k];
            const double xc[3] = {gx.xyz[3 * (size_t)node], gx.xyz[3 * (size_t)node + 1], gx.xyz[3 * (size_t)node + 2]};
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double dn = r < b.elemdim ? b.geo_tab[((size_t)(1 + r) * b.n_geo + k) * b.n_gp + q] : 0.0;
#pragma unroll
                for (int c = 0; c < 3; c++)
                    if (c < b.meshdim) g.J[3 * r + c] += xc[c] * dn;
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            double dn[SLX_MAX_GEO];
#pragma unroll
            for (int k = 0; k < SLX_MAX_GEO; k++)
                dn[k] = (r < b.elemdim && k < b.n_geo) ? b.geo_tab[((size_t)(1 + r) * b.n_geo + k) * b.n_gp + q] : 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < SLX_MAX_GEO; k++) s += x[k][c] * dn[k];
                g.J[3 * r + c] = (c < b.meshdim) ? s : 0.0;
            }
        }
    }
    const double* J = g.J;
#pragma unroll
    for (int i = 0; i < 9; i++) g.inv[i] = 0.0;
    if (b.meshdim == 1) {
        g.det = J[0];
        g.inv[0] = 1.0 / J[0];
    } else if (b.meshdim == 2) {
        g.det = J[0] * J[4] - J[1] * J[3];
        g.inv[0] = J[4] / g.det;
        g.inv[1] = -J[1] / g.det;
        g.inv[3] = -J[3] / g.det;
        g.inv[4] = J[0] / g.det;
    } else {
        g.det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[6] * J[5]) + J[2] * (J[3] * J[7] - J[6] * J[4]);
        const double d = g.det;
        g.inv[0] = (J[4] * J[8] - J[5] * J[7]) / d;
        g.inv[1] = (J[2] * J[7] - J[1] * J[8]) / d;
        g.inv[2] = (J[1] * J[5] - J[2] * J[4]) / d;
        g.inv[3] = (J[5] * J[6] - J[3] * J[8]) / d;
        g.inv[4] = (J[0] * J[8] - J[2] * J[6]) / d;
        g.inv[5] = (J[2] * J[3] - J[0] * J[5]) / d;
        g.inv[6] = (J[3] * J[7] - J[4] * J[6]) / d;
        g.inv[7] = (J[1] * J[6] - J[0] * J[7]) / d;
        g.inv[8] = (J[0] * J[4] - J[1] * J[3]) / d;
    }
}

__device__ inline void geometry(const DevBucket& b, const double* xyz, int e, int q, Geo& g) {
    GeoX gx;
    geo_load(b, xyz, e, gx);
    geo_eval(b, gx, q, g);
}

// Value of field f (space derivative sd) at Gauss point q of element e
// (opfield / rawfield::interpolate: coefficients x form functions, space
// derivatives through the inverse Jacobian).
__device__ inline double field_value(const DevBucket& b, const DevInputs& in, int f, int sd, int e, int q, const Geo& g) {
    const DevField& fl = b.fields[f];
    const double* data = in.field_data[f];
    double v[4] = {0, 0, 0, 0};
    const int nd = sd == 0 ? 1 : 1 + b.elemdim;
    // eight form functions at a time: their address loads, then their value
    // gathers, are issued together (one memory latency per batch instead of
    // two per function); the sum keeps the function order
    constexpr int FB = 8;
    for (int k0 = 0; k0 < fl.nff; k0 += FB) {
        int a[FB];
        double c[FB];
#pragma unroll
        for (int u = 0; u < FB; u++) a[u] = k0 + u < fl.nff ? fl.addr[(size_t)(k0 + u) * b.n_elem + e] : -1;
#pragma unroll
        for (int u = 0; u < FB; u++) c[u] = a[u] >= 0 ? data[a[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < FB; u++) {
            if (a[u] < 0) continue;
#pragma unroll
            for (int d = 0; d < 4; d++)
                if (d < nd) v[d] += c[u] * fl.tab[((size_t)d * fl.nff + k0 + u) * b.n_gp + q];
        }
    }
    if (sd == 0) return v[0];
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (i < b.elemdim) s += pick9(g.inv, 3 * (sd - 1) + i) * v[1 + i];
    return s;
}

// The values of field f's NFF form functions on element e (ok: the function
// has an address), gathered once for all Gauss points a thread takes of the
// element (NFF must equal the bucket's b.fields[f].nff).
template <int NFF>
__device__ inline void field_gather(const DevBucket& b, const DevInputs& in, int f, int e, double (&c)[NFF], bool (&ok)[NFF]) {
    const DevField& fl = b.fields[f];
    const double* data = in.field_data[f];
    int a[NFF];
#pragma unroll
    for (int k = 0; k < NFF; k++) a[k] = fl.addr[(size_t)k * b.n_elem + e];
#pragma unroll
    for (int k = 0; k < NFF; k++) {
        ok[k] = a[k] >= 0;
        c[k] = ok[k] ? data[a[k]] : 0.0;
    }
}

// field_value from the values field_gather took (same terms, same order).
template <int NFF>
__device__ inline double field_value_pre(const DevBucket& b, int f, const double (&c)[NFF], const bool (&ok)[NFF], int sd, int q,
                                         const Geo& g) {
    const DevField& fl = b.fields[f];
    double v[4] = {0, 0, 0, 0};
    const int nd = sd == 0 ? 1 : 1 + b.elemdim;
#pragma unroll
    for (int k = 0; k < NFF; k++) {
        if (!ok[k]) continue;
#pragma unroll
        for (int d = 0; d < 4; d++)
            if (d < nd) v[d] += c[k] * fl.tab[((size_t)d * NFF + k) * b.n_gp + q];
    }
    if (sd == 0) return v[0];
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (i < b.elemdim) s += pick9(g.inv, 3 * (sd - 1) + i) * v[1 + i];
    return s;
}

template <typename T>
__device__ inline T apply_unary(int op, T x) {
    switch (op) {
        case SLX_OP_INV: return (T)1 / x;
        case SLX_OP_ABS: return fabs(x);
        case SLX_OP_EXP: return exp(x);
        case SLX_OP_LOG: return log(x);
        case SLX_OP_SIN: return sin(x);
        case SLX_OP_COS: return cos(x);
        case SLX_OP_TAN: return tan(x);
        case SLX_OP_SQRT: return sqrt(x);
    }
    return x;
}

}  // namespace slx
