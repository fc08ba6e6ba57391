// Symbolic phase on the device: the element-row segments of every free row,
// the CSR pattern with its A/D split, the reduction slots, shared slot blocks
// and the reduction work order (rawmat::process, rawmat.cpp:127-298;
// gentools::findtruefalse, gentools.cpp:2040-2068).  It produces the arrays of
// the host symbolic phase in plan.cpp bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

namespace slx {

// One element-row segment group: the matrix (or rhs) slices of a bucket that
// share a tf address table, stored side by side in the element matrices.
struct SymGroup {
    const int32_t* tf_addr;  // host [n_tf][n_elem] of the group's first slice
    const int8_t* tf_sign;   // host [n_tf][n_elem] of the first slice, or null
    int64_t base;            // element-matrix offset of the group (ascending with the group index)
    int32_t ld, len;         // segment stride and length (sum of the slices' n_dof; 1 for rhs)
    int32_t n_tf, n_elem;
    int32_t slice0, n_slices;  // the group's slices, in segment column order, in the slice table
    int32_t rhs, pad;
};

// One matrix slice of a group.  The tables are the host tables of the plan
// description ([n][n_elem]); the phase uploads each distinct table once and
// transposes what it gathers to element-major order.
struct SymSlice {
    const int32_t* dof_addr;  // [n_dof][n_elem]
    const int8_t* tf_sign;    // [n_tf][n_elem] or null
    const int8_t* dof_sign;   // [n_dof][n_elem] or null
    int32_t n_dof, col_off;   // columns of the slice within the group's segment
    int32_t n_tf, n_elem;
};

struct SymInput {
    int64_t n_dofs = 0;
    const uint8_t* constrained = nullptr;    // host [n_dofs], null: nothing constrained
    std::vector<SymGroup> groups;            // every group, in layout order
    std::vector<SymSlice> slices;
    std::vector<int> mat, rhs;               // group indices in accumulation (fragment rank) order
    bool signed_slots = false;               // folded orientations: signed segment entries
    int n_cu = 256;
};

// A device allocation handed to the caller (hipMalloc'd, caller frees).
struct SymBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct SymOutput {
    int64_t nf = 0, nc = 0, nnzA = 0, nnzD = 0, n_erows = 0, n_coo = 0, n_rhs_erows = 0, n_slots_w = 0;
    int max_erow_len = 1, maxrow = 1, slot_bytes = 1;
    // CSR pattern (Arow/Drow int64 [nf + 1], Acol/Dcol int32), reduction
    // streams in work order (rdesc 4 words per position, erows, slots of the
    // narrow type), rhs segments per row, free and constrained DOF lists
    SymBuf Arow, Drow, Acol, Dcol, rdesc, erows_w, slots_w, rhs_ptr, rhs_erows, free_dofs, cons_dofs;
    std::vector<int64_t> h_Arow, h_Drow;  // host copies of the row pointers
    std::vector<int32_t> h_work_len;      // row lengths (A + D columns) in work order
    bool fallback = false;                // a row exceeds the device kernels' capacity
};

// Runs the phase on `st` and waits for it.  Returns hipSuccess with `err`
// empty on success; on invalid input (an address out of range, two local
// functions of one element row on one DOF, a row longer than the slot
// format) returns hipErrorInvalidValue with the message in `err`.  With
// out.fallback set (and hipSuccess) nothing was produced: the caller runs the
// host phase.
hipError_t symbolic_device(const SymInput& in, SymOutput& out, std::string& err, hipStream_t st);

}  // namespace slx
