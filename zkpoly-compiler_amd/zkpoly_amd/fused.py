"""Fused elementwise kernels: ArithGraph JSON -> HIP library -> operator.

Mirrors (reference paths):
  * FusedOp::gen / gen_fused_kernels: ArithGraph -> kernel source + library
    (core/src/fused_kernels/kernel_gen.rs:47-191,
    compiler/src/transit/type3/lowering/kernel_gen.rs:96-158); here the C++
    generator in libzkp_fused_codegen.so emits HIP and runs hipcc.
  * FusedKernelMeta {name, num_vars, num_mut_vars, lib_path}
    (runtime/src/functions.rs:37-43).
  * FusedKernel::get_fn (core/src/fused_kernels.rs:97-196): vars =
    [stream, inputs...], mut_vars = [arg_buffer, local_buffer, outputs...];
    the PolyPtr arrays are copied into arg_buffer and the kernel is called
    with is_first = true.
  * PipelinedFusedKernel::get_fn (core/src/fused_kernels.rs:198-516): the same
    kernel streamed over host-resident polynomials in divide_parts chunks by
    the native driver in libzkp_fused_pipeline.so (three streams, two device
    buffers per read-only polynomial, three per mutable one).
GraphBuilder writes the serde JSON of common::arith::ArithGraph
(common/src/arith.rs:172-331) so graphs can be made without the compiler.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import hip
from .runtime import (
    PACKAGE_ROOT,
    FuncMeta,
    Function,
    GpuBuffer,
    KernelType,
    Libs,
    PolyPtr,
    RegisteredFunction,
    Scalar,
    ScalarArray,
    unwrap_stream,
)

CODEGEN_LIB = "libzkp_fused_codegen.so"
LIB_NAME = "libfused_kernels.so"
POLYPTR_BYTES = ctypes.sizeof(PolyPtr)


class GraphBuilder:
    """Builds the serde JSON form of ArithGraph<VariableId, ExprId>."""

    def __init__(self, poly_repr: str = "Lagrange", poly_degree: Optional[int] = None):
        assert poly_repr in ("Lagrange", "Coef")
        self.vertices: list = []
        self.inputs: list[int] = []
        self.outputs: list[int] = []
        self.poly_repr = poly_repr
        self.poly_degree = poly_degree
        self._outer = 0

    def _add(self, op) -> int:
        self.vertices.append({"op": op})
        return len(self.vertices) - 1

    def input(self, scalar: bool = False, mutable: bool = False) -> int:
        vid = self._add({"Input": {"outer_id": self._outer, "typ": "Scalar" if scalar else "ScalarArray",
                                   "mutability": "Mut" if mutable else "Const"}})
        self._outer += 1
        self.inputs.append(vid)
        return vid

    def bin(self, group: str, op: str, lhs: int, rhs: int) -> int:
        assert group in ("Pp", "Ss", "Sp")
        return self._add({"Arith": {"Bin": [{group: op}, lhs, rhs]}})

    def neg(self, x: int, scalar: bool = False) -> int:
        return self._add({"Arith": {"Unr": [{"S" if scalar else "P": "Neg"}, x]}})

    def pow(self, x: int, e: int) -> int:
        return self._add({"Arith": {"Unr": [{"P": {"Pow": e}}, x]}})

    def output(self, store: int, scalar: bool = False, in_node: Optional[int] = None) -> int:
        vid = self._add({"Output": {"outer_id": self._outer, "typ": "Scalar" if scalar else "ScalarArray",
                                    "store_node": store, "in_node": in_node}})
        self._outer += 1
        self.outputs.append(vid)
        return vid

    def to_json(self) -> str:
        return json.dumps({"outputs": self.outputs, "inputs": self.inputs, "g": [self.vertices, None],
                           "poly_repr": self.poly_repr, "poly_degree": self.poly_degree})


def graph_ab_gamma_c_minus_z() -> str:
    """out = a*b + gamma*c - z over Lagrange polynomials a, b, c and scalars gamma, z.

    vars order: [a, b, c, gamma, z]; mut_vars: [out].
    """
    g = GraphBuilder("Lagrange")
    a, b, c = g.input(), g.input(), g.input()
    gamma, z = g.input(scalar=True), g.input(scalar=True)
    ab = g.bin("Pp", "Mul", a, b)
    gc = g.bin("Sp", "Mul", gamma, c)
    s = g.bin("Pp", "Add", ab, gc)
    r = g.bin("Sp", "SubBy", z, s)  # s - z
    g.output(r)
    return g.to_json()


def graph_coef_rules() -> str:
    """Coefficient-form scalar-polynomial ops (kernel_gen.rs:503-534).

    vars [p, s]; mut_vars [s + p, s - p, p - s, s * p].
    """
    g = GraphBuilder("Coef")
    p, s = g.input(), g.input(scalar=True)
    for op in ("Add", "Sub", "SubBy", "Mul"):
        g.output(g.bin("Sp", op, s, p))
    return g.to_json()


def graph_inplace_mixed() -> str:
    """In-place output, scalar ops and a scalar output.

    vars [y, s] (x is read through its in-place slot); mut_vars [x', d, t, q] with
    x' = (x + y)^3 * s (overwrites x), d = -(x - y), t = s*s + s (scalar), q = y * t.
    """
    g = GraphBuilder("Lagrange")
    x = g.input(mutable=True)
    y = g.input()
    s = g.input(scalar=True)
    xy = g.bin("Pp", "Add", x, y)
    cube = g.pow(xy, 3)
    new_x = g.bin("Sp", "Mul", s, cube)
    g.output(new_x, in_node=x)
    g.output(g.neg(g.bin("Pp", "Sub", x, y)))
    t = g.bin("Ss", "Add", g.bin("Ss", "Mul", s, s), s)
    g.output(t, scalar=True)
    g.output(g.bin("Sp", "Mul", t, y))
    return g.to_json()


def graph_pipelined_abgcz() -> str:
    """a*b + gamma*c - z with the pipelined argument order, scalars first
    (core/src/fused_kernels.rs:358-386): vars [gamma, z, a, b, c]; mut_vars [out]."""
    g = GraphBuilder("Lagrange")
    gamma, z = g.input(scalar=True), g.input(scalar=True)
    a, b, c = g.input(), g.input(), g.input()
    s = g.bin("Pp", "Add", g.bin("Pp", "Mul", a, b), g.bin("Sp", "Mul", gamma, c))
    g.output(g.bin("Sp", "SubBy", z, s))
    return g.to_json()


def graph_pipelined_mixed() -> str:
    """Pipelined order with a mutable scalar and an in-place polynomial:
    vars [s, y]; mut_vars [t, x', q] with t = s*s + s (scalar), x' = (x + y) * t
    (overwrites x), q = -(x - y)."""
    g = GraphBuilder("Lagrange")
    s = g.input(scalar=True)
    y = g.input()
    x = g.input(mutable=True)
    t = g.bin("Ss", "Add", g.bin("Ss", "Mul", s, s), s)
    g.output(t, scalar=True)
    g.output(g.bin("Sp", "Mul", t, g.bin("Pp", "Add", x, y)), in_node=x)
    g.output(g.neg(g.bin("Pp", "Sub", x, y)))
    return g.to_json()


def graph_pipelined_coef() -> str:
    """Coefficient-form rules in pipelined order: vars [s, p]; mut_vars
    [s + p, s - p, p - s, s * p]; element 0 exists once, in the first chunk."""
    g = GraphBuilder("Coef")
    s, p = g.input(scalar=True), g.input()
    for op in ("Add", "Sub", "SubBy", "Mul"):
        g.output(g.bin("Sp", op, s, p))
    return g.to_json()


def random_graph(seed: int, vertices: int = 600, poly_inputs: int = 12, scalar_inputs: int = 6,
                 inplace: int = 2, poly_outputs: int = 16, scalar_outputs: int = 3, coef: bool = False) -> str:
    """A seeded ArithGraph of about `vertices` vertices with every operation the
    generator accepts (Pp / Sp / Ss add, sub, mul, Sp SubBy, Neg of both kinds,
    Pow), in-place outputs over mutable inputs and scalar outputs, in Lagrange
    or coefficient form -- the size a compiler's fusion pass produces for a
    proof's constraint expressions.  Operands are drawn mostly from recent
    values (long dependent chains) and sometimes from anywhere (long live
    ranges).  vars: the const inputs in order (polynomials, then scalars);
    mut_vars: the in-place outputs, then the polynomial outputs, then the
    scalar outputs."""
    import random

    rng = random.Random(seed)
    g = GraphBuilder("Coef" if coef else "Lagrange")
    mut_in = [g.input(mutable=True) for _ in range(inplace)]
    polys = mut_in + [g.input() for _ in range(poly_inputs)]
    scalars = [g.input(scalar=True) for _ in range(scalar_inputs)]

    unused = []  # computed values nothing reads yet: operands and outputs prefer them

    def pick(pool, kind):
        cands = [u for u in unused if (u in kind_of and kind_of[u] == kind)]
        if cands and rng.random() < 0.5:
            v = cands[-1] if rng.random() < 0.7 else rng.choice(cands)
        elif rng.random() < 0.7:
            v = pool[-1 - min(len(pool) - 1, int(rng.expovariate(0.25)))]
        else:
            v = rng.choice(pool)
        if v in unused:
            unused.remove(v)
        return v

    kind_of = {}

    def made(v, kind):
        kind_of[v] = kind
        unused.append(v)
        return v

    budget = vertices - len(g.vertices) - inplace - poly_outputs - scalar_outputs
    for _ in range(max(0, budget)):
        r = rng.random()
        if r < 0.12:
            scalars.append(made(g.bin("Ss", rng.choice(("Add", "Sub", "Mul")), pick(scalars, "s"), pick(scalars, "s")), "s"))
        elif r < 0.15:
            scalars.append(made(g.neg(pick(scalars, "s"), scalar=True), "s"))
        elif r < 0.50:
            polys.append(made(g.bin("Pp", rng.choice(("Add", "Sub", "Mul", "Mul")), pick(polys, "p"), pick(polys, "p")),
                              "p"))
        elif r < 0.85:
            polys.append(made(g.bin("Sp", rng.choice(("Add", "Sub", "SubBy", "Mul")), pick(scalars, "s"),
                                    pick(polys, "p")), "p"))
        elif r < 0.93:
            polys.append(made(g.neg(pick(polys, "p")), "p"))
        else:
            polys.append(made(g.pow(pick(polys, "p"), rng.randint(2, 9)), "p"))
    computed = polys[len(mut_in) + poly_inputs:]
    for m in mut_in:
        g.output(pick(computed, "p"), in_node=m)
    for _ in range(poly_outputs):
        g.output(pick(computed, "p"))
    for _ in range(scalar_outputs):
        g.output(pick(scalars, "s"), scalar=True)
    return g.to_json()


def demo_kernels() -> list[tuple[str, str]]:
    return [
        (graph_ab_gamma_c_minus_z(), "fused_arith_abgcz"),
        (graph_coef_rules(), "fused_arith_coef_rules"),
        (graph_inplace_mixed(), "fused_arith_inplace_mixed"),
        (graph_pipelined_abgcz(), "fused_arith_pipe_abgcz"),
        (graph_pipelined_mixed(), "fused_arith_pipe_mixed"),
        (graph_pipelined_coef(), "fused_arith_pipe_coef"),
    ]


DEMO_DIR = os.path.join(PACKAGE_ROOT, "lib", "fused_demo")

# Compiler-scale graphs kept prebuilt for the tests (random_graph, Lagrange and
# coefficient form), in a library of their own.
LARGE_DIR = os.path.join(PACKAGE_ROOT, "lib", "fused_large")
LARGE_SEEDS = (101, 202)
LARGE_VERTICES = 800


def large_kernel_name(seed: int, coef: bool) -> str:
    return f"fused_arith_large_{seed}_{'coef' if coef else 'lagrange'}"


def large_kernels() -> list[tuple[str, str]]:
    return [(random_graph(LARGE_SEEDS[0], LARGE_VERTICES, coef=False), large_kernel_name(LARGE_SEEDS[0], False)),
            (random_graph(LARGE_SEEDS[1], LARGE_VERTICES, coef=True), large_kernel_name(LARGE_SEEDS[1], True))]


def _codegen_lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(CODEGEN_LIB)
    if not getattr(lib, "_zkp_typed", False):
        P = ctypes.c_void_p
        lib.zkp_fused_codegen.argtypes = [ctypes.c_char_p, ctypes.c_char_p, P, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p, ctypes.c_size_t]
        lib.zkp_fused_codegen.restype = ctypes.c_int
        lib.zkp_fused_build.argtypes = [P, P, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                        P, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        lib.zkp_fused_build.restype = ctypes.c_int
        lib._zkp_typed = True
    return lib


class CodegenError(ValueError):
    pass


def set_codegen_options(libs: Libs, reg_budget: int = 16, waves_per_eu: int = 2) -> None:
    """Register budget (field elements per element before spilling to
    local_buffer) and waves per SIMD of the kernels generated afterwards."""
    lib = _codegen_lib(libs)
    lib.zkp_fused_set_options.argtypes = [ctypes.c_uint, ctypes.c_uint]
    lib.zkp_fused_set_options.restype = None
    lib.zkp_fused_set_options(reg_budget, waves_per_eu)


def codegen(libs: Libs, graph_json: str, name: str) -> str:
    lib = _codegen_lib(libs)
    err = ctypes.create_string_buffer(1024)
    n = ctypes.c_size_t(0)
    if lib.zkp_fused_codegen(graph_json.encode(), name.encode(), None, 0, ctypes.byref(n), err, 1024) != 0:
        raise CodegenError(err.value.decode())
    buf = ctypes.create_string_buffer(n.value + 1)
    lib.zkp_fused_codegen(graph_json.encode(), name.encode(), buf, n.value + 1, None, err, 1024)
    return buf.value.decode()


def build_library(libs: Libs, kernels: list[tuple[str, str]], out_dir: str) -> str:
    """Generates and compiles [(graph_json, name), ...] into out_dir/libfused_kernels.so."""
    from . import build

    lib = _codegen_lib(libs)
    jsons = (ctypes.c_char_p * len(kernels))(*[k[0].encode() for k in kernels])
    names = (ctypes.c_char_p * len(kernels))(*[k[1].encode() for k in kernels])
    err = ctypes.create_string_buffer(4096)
    os.makedirs(out_dir, exist_ok=True)
    rc = lib.zkp_fused_build(jsons, names, len(kernels), out_dir.encode(), build.hipcc().encode(),
                             os.path.join(PACKAGE_ROOT, "csrc").encode(), None, 0, err, 4096)
    if rc != 0:
        raise CodegenError(err.value.decode())
    return os.path.join(out_dir, LIB_NAME)


@dataclass
class PipelinedMeta:
    """runtime/src/functions.rs:45-50: how many chunks (> 3; lengths must be
    multiples of it) and how many of the vars / mut_vars are scalars."""

    divide_parts: int
    num_scalars: int
    num_mut_scalars: int


@dataclass
class FusedKernelMeta:
    name: str
    num_vars: int
    num_mut_vars: int
    lib_path: str
    pipelined_meta: Optional[PipelinedMeta] = None


def arg_buffer_bytes(meta: FusedKernelMeta) -> int:
    """Size of the argument buffer the compiler allocates (temporary_space.rs:49-56)."""
    return (meta.num_vars + meta.num_mut_vars) * POLYPTR_BYTES


def local_buffer_bytes(meta: FusedKernelMeta) -> int:
    """Size of the kernel's local buffer (the reference's FusedOp::get_temp_buffer_size,
    kernel_gen.rs:136-146): the generated library reports it -- the table of the
    graph's scalar values and the register spill slots of its capped grid; 0
    when the kernel needs none (the buffer may then be empty)."""
    return int(getattr(_load(meta), meta.name + "_local_buffer_bytes")())


def _load(meta: FusedKernelMeta) -> ctypes.CDLL:
    path = os.path.join(meta.lib_path, LIB_NAME)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built")
    lib = ctypes.CDLL(path)
    f = getattr(lib, meta.name)
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_bool, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    fb = getattr(lib, meta.name + "_batched")
    fb.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_ulonglong,
                   ctypes.c_bool, ctypes.c_void_p, ctypes.c_void_p]
    fb.restype = ctypes.c_int
    lb = getattr(lib, meta.name + "_local_buffer_bytes")
    lb.argtypes = []
    lb.restype = ctypes.c_ulonglong
    ar = getattr(lib, meta.name + "_arity")
    ar.argtypes = [ctypes.c_void_p]
    out = (ctypes.c_uint * 2)()
    ar(out)
    if (out[0], out[1]) != (meta.num_vars, meta.num_mut_vars):
        raise ValueError(f"{meta.name}: library arity {tuple(out)} != meta ({meta.num_vars}, {meta.num_mut_vars})")
    return lib


def _ptr_of(v) -> PolyPtr:
    if isinstance(v, ScalarArray):
        return PolyPtr.from_array(v)
    if isinstance(v, Scalar):
        return PolyPtr(v.value, 1, 0, 0, 1)
    raise AssertionError("Only scalars and scalar arrays are supported")


class FusedKernel(RegisteredFunction):
    def __init__(self, libs: Libs, meta: FusedKernelMeta):
        self.meta = meta
        self.lib = _load(meta)
        self.c_func = getattr(self.lib, meta.name)
        self.c_batched = getattr(self.lib, meta.name + "_batched")
        self.local_bytes = int(getattr(self.lib, meta.name + "_local_buffer_bytes")())

    def get_fn(self) -> Function:
        meta, c_func, local_bytes = self.meta, self.c_func, self.local_bytes

        def f(mut_var, var, gpu_mapping=None):
            assert meta.num_vars == len(var) - 1
            assert meta.num_mut_vars == len(mut_var) - 2
            stream = unwrap_stream(var[0])
            arg_buffer, local_buffer = mut_var[0], mut_var[1]
            assert isinstance(arg_buffer, GpuBuffer) and isinstance(local_buffer, GpuBuffer)
            assert arg_buffer.size >= arg_buffer_bytes(meta)
            assert local_buffer.size >= local_bytes, "local buffer too small"
            length = 0
            for v in list(var[1:]) + list(mut_var[2:]):
                if isinstance(v, ScalarArray):
                    if length in (0, 1):
                        length = v.len
                    else:
                        assert length == v.len
                elif length == 0:
                    length = 1
            assert length > 0
            arr = (PolyPtr * (meta.num_vars + meta.num_mut_vars))(
                *[_ptr_of(v) for v in var[1:]], *[_ptr_of(v) for v in mut_var[2:]]
            )
            hip.set_device(stream.get_device())
            hip.check(hip._lib().hipMemcpyAsync(ctypes.c_void_p(arg_buffer.ptr), ctypes.addressof(arr),
                                                ctypes.sizeof(arr), hip.hipMemcpyHostToDevice, stream.raw()),
                      "arg buffer H2D")
            d_vars = arg_buffer.ptr
            d_mut = arg_buffer.ptr + meta.num_vars * POLYPTR_BYTES
            hip.check(c_func(d_vars, d_mut, length, True, local_buffer.ptr, stream.raw()), meta.name)
            # the host array must outlive the async copy
            stream.sync()

        return Function(FuncMeta(meta.name, KernelType.FusedArith, meta), f)


PIPELINE_LIB = "libzkp_fused_pipeline.so"


def _pipeline_lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(PIPELINE_LIB)
    if not getattr(lib, "_zkp_typed", False):
        vp, u = ctypes.c_void_p, ctypes.c_uint
        lib.zkp_fused_pipelined.argtypes = [vp, u, vp, vp, u, vp, vp, u, vp, vp, u, vp, vp, vp, ctypes.c_ulonglong,
                                            u, vp, ctypes.c_size_t, vp, ctypes.c_int]
        lib.zkp_fused_pipelined.restype = ctypes.c_int
        lib._zkp_typed = True
    return lib


def pipelined_arg_buffer_bytes(meta: FusedKernelMeta) -> int:
    """Argument buffer of the pipelined form: two sets of vars, three of
    mut_vars (core/src/fused_kernels.rs:238-241, temporary_space.rs:58-97)."""
    return (2 * meta.num_vars + 3 * meta.num_mut_vars) * POLYPTR_BYTES


class PipelinedFusedKernel(RegisteredFunction):
    """core/src/fused_kernels.rs:198-516.  With p mut scalars, n mut polys,
    q scalars and m polys (all host-resident, polys unsliced):
      mut_var = [arg_buffer, local_buffer, p mut scalars, n mut polys,
                 p scalar GpuBuffers, 3n poly GpuBuffers (set j: [j*n, (j+1)*n))]
      var     = [q scalars, m polys, q scalar GpuBuffers, 2m poly GpuBuffers]
    Poly buffers hold len / divide_parts elements; scalar buffers one."""

    def __init__(self, libs: Libs, meta: FusedKernelMeta):
        assert meta.pipelined_meta is not None
        assert meta.pipelined_meta.divide_parts > 3
        self.kernel = FusedKernel(libs, meta)
        self.meta = meta
        self.driver = _pipeline_lib(libs).zkp_fused_pipelined
        # mut_vars the kernel reads; write-only outputs are not streamed in
        reads = (ctypes.c_ubyte * max(1, meta.num_mut_vars))(*([1] * meta.num_mut_vars))
        f = getattr(self.kernel.lib, meta.name + "_mut_reads", None)
        if f is not None:
            f.argtypes = [ctypes.c_void_p]
            f.restype = None
            f(reads)
        self.mut_reads = [int(x) for x in reads[:meta.num_mut_vars]]

    def get_fn(self) -> Function:
        meta, driver = self.meta, self.driver
        pm = meta.pipelined_meta
        kernel_addr = ctypes.cast(self.kernel.c_func, ctypes.c_void_p).value
        p, q, parts = pm.num_mut_scalars, pm.num_scalars, pm.divide_parts
        poly_reads = self.mut_reads[p:]  # mut_vars are [mut scalars..., mut polys...]

        local_bytes = self.kernel.local_bytes

        def f(mut_var, var, gpu_mapping=None):
            arg_buffer, local_buffer = mut_var[0], mut_var[1]
            assert isinstance(arg_buffer, GpuBuffer) and isinstance(local_buffer, GpuBuffer)
            assert arg_buffer.size >= pipelined_arg_buffer_bytes(meta)
            assert local_buffer.size >= local_bytes, "local buffer too small"
            rest = mut_var[2:]
            assert (len(rest) - 2 * p) % 4 == 0
            assert (len(var) - 2 * q) % 3 == 0
            n = (len(rest) - 2 * p) // 4
            m = (len(var) - 2 * q) // 3
            assert n + p == meta.num_mut_vars
            assert m + q == meta.num_vars
            mut_scalars = [_host_scalar(v) for v in rest[:p]]
            mut_polys = [_host_poly(v) for v in rest[p:p + n]]
            mut_scalar_bufs = [_scalar_buffer(v) for v in rest[p + n:2 * p + n]]
            mut_poly_bufs = list(rest[2 * p + n:])
            scalars = [_host_scalar(v) for v in var[:q]]
            polys = [_host_poly(v) for v in var[q:q + m]]
            scalar_bufs = [_scalar_buffer(v) for v in var[q + m:2 * q + m]]
            poly_bufs = list(var[2 * q + m:])
            length = mut_polys[0].len if n else polys[0].len if m else 1
            assert length % parts == 0
            chunk = length // parts
            for b in mut_poly_bufs + poly_bufs:
                assert isinstance(b, GpuBuffer) and b.size == chunk * 32
            for a in mut_polys + polys:
                assert a.len == length

            def arr(ctype, items):
                return (ctype * max(1, len(items)))(*items)

            c_vp = ctypes.c_void_p
            hs = arr(c_vp, [s.value for s in scalars])
            ds = arr(c_vp, [b.ptr for b in scalar_bufs])
            hms = arr(c_vp, [s.value for s in mut_scalars])
            dms = arr(c_vp, [b.ptr for b in mut_scalar_bufs])
            hp = arr(PolyPtr, [PolyPtr.from_array(a) for a in polys])
            dp = arr(c_vp, [b.ptr for b in poly_bufs])
            hmp = arr(PolyPtr, [PolyPtr.from_array(a) for a in mut_polys])
            dmp = arr(c_vp, [b.ptr for b in mut_poly_bufs])
            reads = arr(ctypes.c_ubyte, poly_reads)
            device = gpu_mapping(0) if gpu_mapping is not None else 0
            hip.check(driver(kernel_addr, q, hs, ds, p, hms, dms, m, hp, dp, n, hmp, dmp, reads, length, parts,
                             arg_buffer.ptr, arg_buffer.size, local_buffer.ptr, device), meta.name + " (pipelined)")

        return Function(FuncMeta(meta.name, KernelType.FusedArith, meta), f)


def _host_scalar(v) -> Scalar:
    assert isinstance(v, Scalar) and v.device == "cpu", "pipelined scalars live in host memory"
    return v


def _host_poly(v) -> ScalarArray:
    assert isinstance(v, ScalarArray) and v.device == "cpu", "pipelined polynomials live in host memory"
    assert v.slice_info is None, "pipelined fused kernel doesn't support slice"
    return v


def _scalar_buffer(v) -> GpuBuffer:
    assert isinstance(v, GpuBuffer) and v.size == 32
    return v
