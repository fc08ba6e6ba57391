// Device-side descriptions shared by the HIP kernels and the host plan code.
#pragma once
#include <stdint.h>
#include "../../../include/slx.h"

#define SLX_MAX_SLICES 12
#define SLX_MAX_FIELDS 4
#define SLX_MAX_REGS 64         // registers of the coefficient interpreter (larger programs run compiled)
#define SLX_MAX_PROG_REGS 256   // registers of a coefficient program
#define SLX_MAX_GEO 8          // geometry nodes held in registers (order-1 geometry)
#define SLX_MAX_GEO_CURVED 27  // order-2 geometry (27-node hexahedra): streamed from memory

namespace slx {

struct DevSlice {
    int32_t n_tf, n_dof, is_rhs, n_terms;
    const slx_term* terms;      // device
    const double* tf_tabq;      // [n_gp][4][tf_pad]  (Gauss point major, zero padded)
    const double* dof_tabq;     // [n_gp][4][dof_pad] or null for right-hand sides
    const float* tf_tabq32;     // FP32 copies of the two tables (FP32 element kernels)
    const float* dof_tabq32;
    int32_t tf_pad, dof_pad;
    int64_t ke_base;            // Ke entry (e, i, j) at ke_base + (e*n_tf + i)*ke_ld + j
    int64_t ke_ld;              // element-row segment length (1 for right-hand sides)
};

struct DevField {
    int32_t nff;
    const int32_t* addr;        // [nff][n_elem]
    const double* tab;          // [4][nff][n_gp]
};

struct DevBucket {
    int32_t type, n_elem, n_geo, n_gp, elemdim, meshdim;
    int32_t n_instr, n_regs, n_coefs, q_inv, n_slices, n_fields;
    const int32_t* geo_nodes;   // [n_elem][n_geo]
    const int32_t* elem_ids;    // [n_elem]
    const double* gp_w;         // [n_gp]
    const double* geo_tab;      // [4][n_geo][n_gp] order-1 Lagrange
    const slx_instr* prog;
    const DevSlice* dslices;    // device copy of the slice descriptions
    DevSlice slices[SLX_MAX_SLICES];
    DevField fields[SLX_MAX_FIELDS];
};

struct DevInputs {
    const double* xyz;          // [n_nodes][3]
    const double* params;       // [n_params][param_len]
    int64_t param_len;
    const double* consts;
    const double* field_data[SLX_MAX_FIELDS];
};

// Reduction maps of the cached CSR pattern.
struct DevReduce {
    int64_t nrows;                  // free rows
    // per work position (rows by decreasing length), 4 words:
    //   Arow | lenA << 48,  Drow | lenD << 48,  first segment | n_segments << 48,  first slot
    const uint64_t* rdesc;
    const uint64_t* erows;          // packed (len << 48) | ke offset; a row's segments consecutive, rows in work order
    const void* slots;              // per segment entry (blocks padded to 4): A position, lenA + D position, or all ones
    int32_t slot_bytes;             // 1, 2 or 4
    int32_t slot_signed;            // 1: the top slot bit negates the entry (folded orientations)
    int32_t n_cu;                   // compute units
    int32_t max_row_len;            // max (lenA + lenD)
    const int64_t* rhs_erow_ptr;    // [nrows+1]
    const uint64_t* rhs_erows;      // packed (1 << 48) | ke offset, bit 63: negate
    const int64_t* free_dofs;       // [nrows] global DOF of each free row
};

// Constant-address-space views: uniform loads through them become scalar loads.
#define SLX_CONST __attribute__((address_space(4)))

}  // namespace slx
