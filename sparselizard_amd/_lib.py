"""ctypes binding of the slx C ABI (include/slx.h).

The library is built in-tree (sparselizard_amd/lib/libslx.so).  Loading
fails loudly when it is missing: there is no CPU fallback for the assembly
path.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libslx.so")

SLX_OK = 0
PREC_FP64 = 0
PREC_F16_SPLIT = 1
PREC_BF16_SPLIT = 2
PREC_FP32 = 3
PREC_FP32_MFMA = 4
PRECISIONS = {"fp64": PREC_FP64, "f16split": PREC_F16_SPLIT, "bf16split": PREC_BF16_SPLIT, "fp32": PREC_FP32,
              "fp32mfma": PREC_FP32_MFMA}


class NumericDesc(C.Structure):
    _fields_ = [
        ("node_xyz", C.POINTER(C.c_double)),
        ("n_params", C.c_int32),
        ("param_len", C.c_int64),
        ("params", C.POINTER(C.c_double)),
        ("field_data", C.POINTER(C.POINTER(C.c_double))),
        ("field_len", C.POINTER(C.c_int64)),
        ("precision", C.c_int32),
    ]


class Instr(C.Structure):
    _fields_ = [("op", C.c_int16), ("dst", C.c_int16), ("a", C.c_int16), ("b", C.c_int16), ("imm", C.c_int32)]


class Term(C.Structure):
    _fields_ = [("tf_deriv", C.c_int32), ("dof_deriv", C.c_int32), ("coef", C.c_int32)]


class SliceDesc(C.Structure):
    _fields_ = [("frag_rank", C.c_int32), ("is_rhs", C.c_int32), ("n_tf", C.c_int32), ("n_dof", C.c_int32),
                ("tf_addr", C.POINTER(C.c_int32)), ("dof_addr", C.POINTER(C.c_int32)), ("n_terms", C.c_int32),
                ("terms", C.POINTER(Term)), ("tf_table", C.c_int32), ("dof_table", C.c_int32),
                ("tf_sign", C.POINTER(C.c_int8)), ("dof_sign", C.POINTER(C.c_int8))]


class BucketDesc(C.Structure):
    _fields_ = [("elem_type", C.c_int32), ("n_elem", C.c_int32), ("n_geo", C.c_int32),
                ("geo_nodes", C.POINTER(C.c_int32)), ("elem_ids", C.POINTER(C.c_int32)), ("n_gp", C.c_int32),
                ("gp_coords", C.POINTER(C.c_double)), ("gp_weights", C.POINTER(C.c_double)), ("n_tables", C.c_int32),
                ("table_nff", C.POINTER(C.c_int32)), ("tables", C.POINTER(C.POINTER(C.c_double))),
                ("n_slices", C.c_int32), ("slices", C.POINTER(SliceDesc)), ("n_instr", C.c_int32),
                ("prog", C.POINTER(Instr)), ("n_regs", C.c_int32), ("n_coefs", C.c_int32), ("q_invariant", C.c_int32)]


class PlanDesc(C.Structure):
    _fields_ = [("mesh_dim", C.c_int32), ("n_nodes", C.c_int64), ("n_dofs", C.c_int64),
                ("constrained", C.POINTER(C.c_uint8)), ("n_consts", C.c_int32), ("consts", C.POINTER(C.c_double)),
                ("n_buckets", C.c_int32), ("buckets", C.POINTER(BucketDesc)), ("n_fields", C.c_int32),
                ("field_nff", C.POINTER(C.c_int32)), ("field_addr", C.POINTER(C.POINTER(C.c_int32))),
                ("field_tables", C.POINTER(C.POINTER(C.c_double))), ("has_rhs", C.c_int32), ("has_matrix", C.c_int32),
                ("dirichlet", C.POINTER(C.c_double))]


PlanDescHead = PlanDesc


class SlxError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libslx.so (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SlxError(
            "libslx.so is not built (%s); run __graft_entry__.build() or `make -C sparselizard_amd`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32, i64 = C.c_int32, C.c_int64
    p64, p32, pd = C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    sig = {
        "slx_last_error": (C.c_char_p, []),
        "slx_create": (i32, [C.c_int, C.POINTER(vp)]),
        "slx_destroy": (i32, [vp]),
        "slx_plan_create": (i32, [vp, vp, C.POINTER(vp)]),
        "slx_plan_create_ex": (i32, [vp, vp, i32, C.POINTER(vp)]),
        "slx_plan_symbolic_info": (i32, [vp, p32, pd]),
        "slx_plan_symbolic_array": (i32, [vp, i32, vp, i64, p64]),
        "slx_plan_shape": (i32, [vp, p64, p64, p64, p64]),
        "slx_plan_stats": (i32, [vp, p64, p64, p64]),
        "slx_assemble": (i32, [vp, C.POINTER(NumericDesc)]),
        "slx_assemble_device": (i32, [vp, vp, vp, i64, vp, i32, vp]),
        "slx_plan_sync": (i32, [vp]),
        "slx_assemble_async": (i32, [vp, C.POINTER(NumericDesc), vp]),
        "slx_plan_dof_maps": (i32, [vp, p64, p64]),
        "slx_plan_profile": (i32, [vp, vp, vp, i64, vp, i32, i32, pd]),
        "slx_csr_device": (i32, [vp] + [C.POINTER(vp)] * 7),
        "slx_csr_copyout": (i32, [vp, p64, p32, pd, p64, p32, pd, pd]),
        "slx_plan_destroy": (i32, [vp]),
        "slx_plan_kernels": (i32, [vp, p32, p32]),
        "slx_plan_kernel_counts": (i32, [vp, i32, p32, p32]),
        "slx_plan_tile_count": (i32, [vp, i32, p32]),
        "slx_plan_fused_info": (i32, [vp, i32, p32, p64]),
        "slx_plan_solve": (i32, [vp, i32, C.c_double, i32, i32, pd, vp, p32, p32, pd, p32]),
        "slx_plan_solve_noalloc": (i32, [vp, i32, C.c_double, i32, i32, pd, vp, p32, p32, pd, p32]),
        "slx_plan_eliminate": (i32, [vp, pd, pd]),
        "slx_plan_coef_mode": (i32, [vp, i32, p32]),
        "slx_plan_matvec": (i32, [vp, pd, pd]),
        "slx_selftest_spmv_stream": (i32, [vp, i64, p64, p32, pd, pd, pd, pd]),
        "slx_plan_element_matrices": (i32, [vp, pd, i64]),
        "slx_plan_coefficients": (i32, [vp, pd, i64, p64]),
        "slx_plan_reduction_maps": (i32, [vp, p64, p64, C.POINTER(C.c_uint64), p64, p32, p64, C.POINTER(C.c_uint64), p64, p64]),
        "slx_problem_create": (i32, [vp, C.c_char_p, C.c_char_p, C.POINTER(vp)]),
        "slx_problem_destroy": (i32, [vp]),
        "slx_problem_plan": (i32, [vp, C.POINTER(vp)]),
        "slx_problem_numeric": (i32, [vp, C.POINTER(NumericDesc)]),
        "slx_problem_generate": (i32, [vp, i32]),
        "slx_problem_info": (i32, [vp, p64, p64, p64, p64]),
        "slx_problem_desc": (i32, [vp, C.POINTER(vp)]),
        "slx_problem_setdata": (i32, [vp, i32, pd, i64]),
        "slx_problem_setdata_vec": (i32, [vp, pd, i64]),
        "slx_plan_graph_capture": (i32, [vp, vp, vp, i64, vp, i32]),
        "slx_plan_graph_launch": (i32, [vp, vp]),
        "slx_gauss_rule": (i32, [i32, i32, i32, p32, pd, pd]),
        "slx_selftest_mfma_layout": (i32, [vp, C.POINTER(C.c_float)]),
        "slx_selftest_mfma_layout_kind": (i32, [vp, i32, C.POINTER(C.c_float)]),
        "slx_h1_count": (i32, [i32, i32, p32]),
        "slx_h1_tables": (i32, [i32, i32, i32, i32, pd, pd]),
        "slx_geometry_tables": (i32, [i32, i32, i32, pd, pd]),
        "slx_hcurl_count": (i32, [i32, i32, p32]),
        "slx_hcurl_tables": (i32, [i32, i32, i32, i32, i32, pd, pd]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = [
    "slx_last_error", "slx_create", "slx_destroy", "slx_plan_create", "slx_plan_create_ex", "slx_plan_symbolic_info",
    "slx_plan_symbolic_array", "slx_plan_shape", "slx_assemble",
    "slx_assemble_device", "slx_assemble_async", "slx_plan_dof_maps", "slx_plan_sync", "slx_plan_profile", "slx_csr_device", "slx_csr_copyout", "slx_plan_stats",
    "slx_plan_destroy", "slx_plan_kernels", "slx_plan_kernel_counts", "slx_plan_tile_count", "slx_plan_fused_info", "slx_plan_solve", "slx_plan_solve_noalloc", "slx_plan_eliminate", "slx_plan_matvec", "slx_plan_coef_mode", "slx_plan_reduction_maps", "slx_plan_element_matrices", "slx_plan_coefficients", "slx_problem_create", "slx_problem_destroy", "slx_problem_plan", "slx_problem_numeric",
    "slx_problem_generate", "slx_problem_info", "slx_problem_desc", "slx_problem_setdata", "slx_problem_setdata_vec",
    "slx_plan_graph_capture", "slx_plan_graph_launch",
    "slx_gauss_rule", "slx_h1_count", "slx_h1_tables", "slx_geometry_tables", "slx_hcurl_count", "slx_hcurl_tables", "slx_selftest_mfma_layout", "slx_selftest_mfma_layout_kind",
    "slx_selftest_spmv_stream",
]


def mfma_layout_probe(ctx, kind=0):
    """16x16 result of the MFMA layout probe: kind 0 f16 16x16x32
    (slx_selftest_mfma_layout), 1 bf16 16x16x32, 2 f32 16x16x4, 3 f64 16x16x4
    (slx_selftest_mfma_layout_kind)."""
    out = np.zeros(256, np.float32)
    if kind == 0:
        check(lib().slx_selftest_mfma_layout(ctx.h, out.ctypes.data_as(C.POINTER(C.c_float))))
    else:
        check(lib().slx_selftest_mfma_layout_kind(ctx.h, kind, out.ctypes.data_as(C.POINTER(C.c_float))))
    return out.reshape(16, 16)


def spmv_stream_probe(ctx, rowp, col, val, x):
    """The solve's streaming FP32 CSR product on a host CSR: (y, x.y)."""
    rowp = np.ascontiguousarray(rowp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    n = len(rowp) - 1
    y = np.zeros(n, np.float64)
    xy = C.c_double()
    check(lib().slx_selftest_spmv_stream(ctx.h, n, _p(rowp, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double),
                                         _p(x, C.c_double), _p(y, C.c_double), C.byref(xy)))
    return y, xy.value


def gauss_rule(elem_type, order):
    """(coords [n,3], weights [n]) of gausspoints(type, order)."""
    n = C.c_int32()
    check(lib().slx_gauss_rule(elem_type, order, 0, C.byref(n), None, None))
    xyz = np.zeros((n.value, 3), np.float64)
    w = np.zeros(n.value, np.float64)
    check(lib().slx_gauss_rule(elem_type, order, n.value, C.byref(n), _p(xyz, C.c_double), _p(w, C.c_double)))
    return xyz, w


def h1_count(elem_type, order):
    n = C.c_int32()
    check(lib().slx_h1_count(elem_type, order, C.byref(n)))
    return n.value


def geometry_tables(elem_type, n_geo, pts):
    """Geometry functions and reference derivatives: array [4, n_geo, npts]."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((4, n_geo, len(pts)), np.float64)
    check(lib().slx_geometry_tables(elem_type, n_geo, len(pts), _p(pts, C.c_double), _p(out, C.c_double)))
    return out


def h1_tables(elem_type, order, pts, total_orientation=0):
    """Values and reference derivatives: array [4, nff, npts]."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    nff = h1_count(elem_type, order)
    out = np.zeros((4, nff, len(pts)), np.float64)
    check(lib().slx_h1_tables(elem_type, order, total_orientation, len(pts), _p(pts, C.c_double), _p(out, C.c_double)))
    return out


def hcurl_count(elem_type, order):
    n = C.c_int32()
    check(lib().slx_hcurl_count(elem_type, order, C.byref(n)))
    return n.value


HCURL_VALUE, HCURL_CURL = 1, 2


def hcurl_tables(elem_type, order, pts, family=HCURL_VALUE, total_orientation=0):
    """H(curl) values (family 1) or curls (family 2) in reference coordinates:
    array [4, nff, npts], slot 0 zero, slots 1..3 the ki, eta, phi components."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    nff = hcurl_count(elem_type, order)
    out = np.zeros((4, nff, len(pts)), np.float64)
    check(lib().slx_hcurl_tables(elem_type, order, total_orientation, family, len(pts), _p(pts, C.c_double),
                                 _p(out, C.c_double)))
    return out


def check(status):
    if status != SLX_OK:
        raise SlxError("slx error %d: %s" % (status, lib().slx_last_error().decode()))


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


# Live handles, destroyed by _shutdown while the interpreter and the HIP
# runtime are both still intact: Python runs atexit handlers before it tears
# down modules and before the process runs the runtime's own exit handlers.
# Handles that are still alive after that (objects freed late in interpreter
# finalization, e.g. through a failed test's traceback) are left to the
# process exit instead of being destroyed into a runtime that may be gone.
_live = {"plan": weakref.WeakSet(), "problem": weakref.WeakSet(), "ctx": weakref.WeakSet()}
_shut_down = False


def _shutdown():
    global _shut_down
    if _shut_down:
        return
    # plans and problems hold device memory and a context reference: first
    for kind in ("plan", "problem", "ctx"):
        for obj in list(_live[kind]):
            try:
                obj.close()
            except Exception:
                pass
    _shut_down = True


atexit.register(_shutdown)


class Context:
    """One device + stream (slx_create)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().slx_create(device, C.byref(self.h)))
        _live["ctx"].add(self)

    def close(self):
        if self.h:
            lib().slx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _shut_down:
            return
        try:
            self.close()
        except Exception:
            pass


class CSR:
    """Host copy of an assembled A/D block pair and the rhs."""

    def __init__(self, Arow, Acol, Aval, Drow, Dcol, Dval, rhs):
        self.Arow, self.Acol, self.Aval = Arow, Acol, Aval
        self.Drow, self.Dcol, self.Dval = Drow, Dcol, Dval
        self.rhs = rhs

    @property
    def nfree(self):
        return len(self.Arow) - 1


class Problem:
    """A synthetic workload built by the C++ host layer (slx_problem_*)."""

    def __init__(self, ctx, kind, spec=""):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().slx_problem_create(ctx.h if ctx else None, kind.encode(), spec.encode(), C.byref(self.h)))
        self._plan = None
        _live["problem"].add(self)

    def info(self):
        v = [C.c_int64() for _ in range(4)]
        check(lib().slx_problem_info(self.h, *[C.byref(x) for x in v]))
        return dict(ndofs=v[0].value, nfree=v[1].value, nelem=v[2].value, nbuckets=v[3].value)

    def desc_ptr(self):
        d = C.c_void_p()
        check(lib().slx_problem_desc(self.h, C.byref(d)))
        return d

    def desc(self):
        """The plan description (a ctypes view valid while the problem lives)."""
        return C.cast(self.desc_ptr(), C.POINTER(PlanDesc)).contents

    def n_nodes(self):
        return self.desc().n_nodes

    def numeric(self):
        nd = NumericDesc()
        check(lib().slx_problem_numeric(self.h, C.byref(nd)))
        return nd

    def plan(self):
        if self._plan is None:
            p = C.c_void_p()
            check(lib().slx_problem_plan(self.h, C.byref(p)))
            self._plan = Plan(p, owned=False)
        return self._plan

    def generate(self, precision="fp64"):
        check(lib().slx_problem_generate(self.h, PRECISIONS[precision] if isinstance(precision, str) else precision))
        return self.plan()

    def setdata(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().slx_problem_setdata(self.h, field, _p(v, C.c_double), len(v)))

    def setdata_vec(self, x):
        """sl::setdata(vec): field data from a solution vector in DOF numbering."""
        v = np.ascontiguousarray(x, dtype=np.float64)
        check(lib().slx_problem_setdata_vec(self.h, _p(v, C.c_double), len(v)))

    def field_data(self):
        """Host views of the data of the fields read by the coefficient programs."""
        nd = self.numeric()
        d = self.desc()
        out = []
        for k in range(d.n_fields):
            n = int(nd.field_len[k])
            out.append(np.ctypeslib.as_array(nd.field_data[k], shape=(n,)) if n else np.zeros(0))
        return out

    def close(self):
        if self.h:
            if self._plan is not None:
                self._plan.h = None  # owned by the problem: destroyed with it
            lib().slx_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _shut_down:
            return
        try:
            self.close()
        except Exception:
            pass


PLAN_HOST_SYMBOLIC = 1
SYMBOLIC_ARRAYS = ("rdesc", "erows", "slots", "rhs_ptr", "rhs_erows", "free_dofs", "cons_dofs", "row_blocks", "classes",
                   "scalars")


class Plan:
    def __init__(self, h, owned=True):
        self.h = h
        self.owned = owned
        if owned:
            _live["plan"].add(self)

    @classmethod
    def create(cls, ctx, problem, flags=0):
        """A plan of the problem's description (slx_plan_create_ex); the
        problem must outlive it.  flags: PLAN_HOST_SYMBOLIC."""
        p = C.c_void_p()
        check(lib().slx_plan_create_ex(ctx.h if ctx else None, problem.desc_ptr(), flags, C.byref(p)))
        return cls(p, owned=True)

    def symbolic_info(self):
        """(symbolic phase ran on the device, its wall time in seconds)."""
        a, t = C.c_int32(), C.c_double()
        check(lib().slx_plan_symbolic_info(self.h, C.byref(a), C.byref(t)))
        return bool(a.value), t.value

    def symbolic_array(self, name):
        """Raw bytes (numpy uint8) of one device array of the symbolic phase."""
        which = SYMBOLIC_ARRAYS.index(name)
        n = C.c_int64()
        check(lib().slx_plan_symbolic_array(self.h, which, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint8)
        check(lib().slx_plan_symbolic_array(self.h, which, out.ctypes.data_as(C.c_void_p), out.size, C.byref(n)))
        return out[:n.value]

    def shape(self):
        v = [C.c_int64() for _ in range(4)]
        check(lib().slx_plan_shape(self.h, *[C.byref(x) for x in v]))
        return dict(nfree=v[0].value, ncons=v[1].value, nnzA=v[2].value, nnzD=v[3].value)

    def stats(self):
        v = [C.c_int64() for _ in range(3)]
        check(lib().slx_plan_stats(self.h, *[C.byref(x) for x in v]))
        return dict(n_elem_rows=v[0].value, n_ke=v[1].value, n_coo=v[2].value)

    def kernels(self):
        a, b = C.c_int32(), C.c_int32()
        check(lib().slx_plan_kernels(self.h, C.byref(a), C.byref(b)))
        return dict(mfma_buckets=a.value, buckets=b.value)

    def kernel_counts(self, precision):
        """(batched MFMA buckets, factorized MFMA slices) of a precision mode."""
        a, b = C.c_int32(), C.c_int32()
        check(lib().slx_plan_kernel_counts(self.h, PRECISIONS[precision] if isinstance(precision, str) else precision,
                                           C.byref(a), C.byref(b)))
        return dict(batched=a.value, factorized=b.value)

    def tile_count(self, precision):
        """Matrix slices on the one-tile-per-element MFMA kernel in a precision mode."""
        a = C.c_int32()
        check(lib().slx_plan_tile_count(self.h, PRECISIONS[precision] if isinstance(precision, str) else precision, C.byref(a)))
        return a.value

    def fused_info(self, precision):
        """(slices, work items) of the mode's fused launch; (0, 0) without one."""
        a, b = C.c_int32(), C.c_int64()
        check(lib().slx_plan_fused_info(self.h, PRECISIONS[precision] if isinstance(precision, str) else precision,
                                        C.byref(a), C.byref(b)))
        return a.value, b.value

    def element_matrices(self):
        n_ke = self.stats()["n_ke"]
        out = np.zeros(max(n_ke, 1), np.float64)
        check(lib().slx_plan_element_matrices(self.h, _p(out, C.c_double), out.size))
        return out[:n_ke]

    def coefficients(self):
        """Coefficient table of the last bucket of the last assembly (FP64 copy)."""
        n = C.c_int64()
        check(lib().slx_plan_coefficients(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.float64)
        check(lib().slx_plan_coefficients(self.h, _p(out, C.c_double), out.size, C.byref(n)))
        return out[:n.value]

    def ke_layout(self):
        """(ke_base, ke_ld) of every (bucket, slice)."""
        sz = np.zeros(4, np.int64)
        check(lib().slx_plan_reduction_maps(self.h, _p(sz, C.c_int64), None, None, None, None, None, None, None, None))
        kb, kl = np.zeros(sz[3], np.int64), np.zeros(sz[3], np.int64)
        check(lib().slx_plan_reduction_maps(self.h, _p(sz, C.c_int64), None, None, None, None, None, None,
                                            _p(kb, C.c_int64), _p(kl, C.c_int64)))
        return kb, kl

    def reduction_maps(self):
        """Reduction maps and element-matrix layout of a pattern-only plan."""
        sz = np.zeros(4, np.int64)
        check(lib().slx_plan_reduction_maps(self.h, _p(sz, C.c_int64), None, None, None, None, None, None, None, None))
        nf = self.shape()["nfree"]
        out = dict(erow_ptr=np.zeros(nf + 1, np.int64), erows=np.zeros(sz[0], np.uint64),
                   slot_ptr=np.zeros(nf + 1, np.int64), slots=np.zeros(sz[1], np.int32),
                   rhs_ptr=np.zeros(nf + 1, np.int64), rhs_erows=np.zeros(sz[2], np.uint64),
                   ke_base=np.zeros(sz[3], np.int64), ke_ld=np.zeros(sz[3], np.int64))
        check(lib().slx_plan_reduction_maps(self.h, _p(sz, C.c_int64), _p(out["erow_ptr"], C.c_int64),
                                            _p(out["erows"], C.c_uint64), _p(out["slot_ptr"], C.c_int64),
                                            _p(out["slots"], C.c_int32), _p(out["rhs_ptr"], C.c_int64),
                                            _p(out["rhs_erows"], C.c_uint64), _p(out["ke_base"], C.c_int64),
                                            _p(out["ke_ld"], C.c_int64)))
        return out

    def assemble(self, numeric, precision):
        numeric.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
        check(lib().slx_assemble(self.h, C.byref(numeric)))

    def dof_maps(self):
        """mat::getainds / getdinds: DOF numbers of A's rows (free) and D's columns (constrained)."""
        s = self.shape()
        a = np.zeros(max(s["nfree"], 1), np.int64)
        d = np.zeros(max(s["ncons"], 1), np.int64)
        check(lib().slx_plan_dof_maps(self.h, _p(a, C.c_int64), _p(d, C.c_int64)))
        return a[:s["nfree"]], d[:s["ncons"]]

    def assemble_async(self, numeric, precision, stream=None):
        numeric.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
        check(lib().slx_assemble_async(self.h, C.byref(numeric), stream))

    def assemble_device(self, d_xyz, d_params, param_len, d_fields, precision, stream=None):
        check(lib().slx_assemble_device(self.h, d_xyz, d_params, param_len, d_fields,
                                        PRECISIONS[precision] if isinstance(precision, str) else precision, stream))

    def graph_capture(self, d_xyz, d_params, param_len, d_fields, precision):
        check(lib().slx_plan_graph_capture(self.h, d_xyz, d_params, param_len, d_fields,
                                           PRECISIONS[precision] if isinstance(precision, str) else precision))

    def graph_launch(self, stream=None):
        check(lib().slx_plan_graph_launch(self.h, stream))

    def profile(self, d_xyz, d_params, param_len, d_fields, precision, iters=5):
        """Average ms per numeric phase: coef, mfma, valu, reduce, rhs (HIP events per launch)."""
        ms = np.zeros(5, np.float64)
        check(lib().slx_plan_profile(self.h, d_xyz, d_params, param_len, d_fields,
                                     PRECISIONS[precision] if isinstance(precision, str) else precision, iters,
                                     _p(ms, C.c_double)))
        return {"coef": float(ms[0]), "mfma": float(ms[1]), "valu": float(ms[2]), "reduce": float(ms[3]),
                "rhs": float(ms[4])}

    def sync(self):
        check(lib().slx_plan_sync(self.h))

    def csr(self, ndofs):
        s = self.shape()
        nf = s["nfree"]
        Arow = np.zeros(nf + 1, np.int64)
        Drow = np.zeros(nf + 1, np.int64)
        Acol = np.zeros(max(s["nnzA"], 1), np.int32)
        Dcol = np.zeros(max(s["nnzD"], 1), np.int32)
        Aval = np.zeros(max(s["nnzA"], 1), np.float64)
        Dval = np.zeros(max(s["nnzD"], 1), np.float64)
        rhs = np.zeros(max(ndofs, 1), np.float64)
        check(lib().slx_csr_copyout(self.h, _p(Arow, C.c_int64), _p(Acol, C.c_int32), _p(Aval, C.c_double),
                                    _p(Drow, C.c_int64), _p(Dcol, C.c_int32), _p(Dval, C.c_double),
                                    _p(rhs, C.c_double)))
        return CSR(Arow, Acol[:s["nnzA"]], Aval[:s["nnzA"]], Drow, Dcol[:s["nnzD"]], Dval[:s["nnzD"]], rhs[:ndofs])

    SOLVE_METHODS = {"auto": 0, "cg": 1, "gmres": 2}

    def solve(self, ndofs, rel_tol=1e-10, max_outer=20, max_inner=2000, method="auto"):
        """sl::solve(A, b) on the last assembled system (sl.cpp:1565-1627):
        solution in DOF numbering, constrained entries = Dirichlet values.
        method: "cg" (SPD A), "gmres" (any nonsingular A) or "auto" (symmetry
        check).  Returns (x, info) with the refinement count, the inner
        iteration count, the final relative FP64 residual of b_A - D x_D - A x
        and the method used."""
        return self._solve(lib().slx_plan_solve, ndofs, rel_tol, max_outer, max_inner, method)

    def solve_noalloc(self, ndofs, rel_tol=1e-10, max_outer=20, max_inner=2000, method="auto"):
        """The same solve with its workspace in the library's fixed per-GPU
        array (slx_plan_solve_noalloc): no device allocation."""
        return self._solve(lib().slx_plan_solve_noalloc, ndofs, rel_tol, max_outer, max_inner, method)

    def _solve(self, fn, ndofs, rel_tol, max_outer, max_inner, method):
        x = np.zeros(max(ndofs, 1), np.float64)
        o, i, mu = C.c_int32(), C.c_int32(), C.c_int32()
        r = C.c_double()
        check(fn(self.h, self.SOLVE_METHODS[method], rel_tol, max_outer, max_inner, _p(x, C.c_double), None,
                 C.byref(o), C.byref(i), C.byref(r), C.byref(mu)))
        names = {v: k for k, v in self.SOLVE_METHODS.items()}
        return x[:ndofs], dict(outer=o.value, inner=i.value, rel_residual=r.value, method=names.get(mu.value, "?"))

    COEF_MODES = {"auto": 0, "interpreter": 1, "compiled": 2}

    def coef_mode(self, mode):
        """Coefficient programs: "auto" (compiled for buckets of >= 4096 elements),
        "interpreter" or "compiled"; returns the compiled kernels in use."""
        n = C.c_int32()
        check(lib().slx_plan_coef_mode(self.h, self.COEF_MODES[mode], C.byref(n)))
        return n.value

    def eliminate(self, nfree, b=None):
        """mat::eliminate(b) (mat.cpp:78-103): b_A - D b_D (b None: the assembled rhs)."""
        out = np.zeros(max(nfree, 1), np.float64)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        check(lib().slx_plan_eliminate(self.h, None if bb is None else _p(bb, C.c_double), _p(out, C.c_double)))
        return out[:nfree]

    def matvec(self, x):
        """mat * vec (mat.cpp:182-205): y_A = A x_A + D x_D, constrained entries of y = those of x."""
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(max(len(xx), 1), np.float64)
        check(lib().slx_plan_matvec(self.h, _p(xx, C.c_double), _p(y, C.c_double)))
        return y[:len(xx)]

    def device_ptrs(self):
        ptrs = [C.c_void_p() for _ in range(7)]
        check(lib().slx_csr_device(self.h, *[C.byref(x) for x in ptrs]))
        return dict(zip(["Arow", "Acol", "Aval", "Drow", "Dcol", "Dval", "rhs"], [p.value for p in ptrs]))

    def close(self):
        if self.owned and self.h:
            lib().slx_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        if _shut_down:
            return
        try:
            self.close()
        except Exception:
            pass
