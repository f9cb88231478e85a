"""Fused kernels at the size a compiler's fusion pass produces: seeded random
ArithGraphs of 800 vertices (fused.random_graph) with every operation the
generator accepts, in-place and scalar outputs, Lagrange and coefficient
form, checked element by element against an interpreter of the graph over
the oracle's field operations.

The generated kernels hold at most a fixed number of field elements in
registers per element and spill the rest to local_buffer, and keep the
graph's many scalar values in a table that a one-thread prologue fills, so
these graphs run with several waves per SIMD and no compiler scratch; the
CPU tests check the generator's plan for them (spill slots, table, buffer
size), the GPU tests the results.  The reference has no fused-kernel test of
its own (its kernels run only inside compiled programs).
"""
import json

import numpy as np
import pytest

import oracle as orc
from helpers import device_array, gpu2cpu, cpu2gpu
from zkpoly_amd import fused, hip
from zkpoly_amd.field import random_elements
from zkpoly_amd.runtime import GpuBuffer, Libs, Scalar, ScalarArray

FR = orc.FR
CASES = [(fused.LARGE_SEEDS[0], False), (fused.LARGE_SEEDS[1], True)]


def interpret(graph, polys, scalars, n):
    """Values of every vertex: (n, 4) arrays for polynomials, (4,) for scalars.
    polys / scalars: input vertex id -> host values."""
    V = graph["g"][0]
    val = {}
    zero = np.zeros((n, 4), dtype=np.uint64)

    def bcast(x):
        return np.broadcast_to(x, (n, 4)).copy()

    def get(i):
        if i in val:
            return val[i]
        op = V[i]["op"]
        kind = next(iter(op))
        if kind == "Input":
            val[i] = polys[i] if i in polys else scalars[i]
            return val[i]
        a = op["Arith"]
        if "Unr" in a:
            (grp, u), arg = next(iter(a["Unr"][0].items())), a["Unr"][1]
            x = get(arg)
            if u == "Neg":
                if grp == "P":
                    val[i] = orc.vec("orc_vec_sub", FR, zero, x)
                else:
                    val[i] = orc.vec("orc_vec_sub", FR, zero[:1], x.reshape(1, 4))[0]
            else:
                val[i] = orc.vec_pow(FR, x, u["Pow"])
            return val[i]
        (grp, b), l, r = next(iter(a["Bin"][0].items())), a["Bin"][1], a["Bin"][2]
        x, y = get(l), get(r)
        name = {"Add": "orc_vec_add", "Sub": "orc_vec_sub", "Mul": "orc_vec_mul", "SubBy": "orc_vec_sub"}[b]
        if grp == "Ss":
            val[i] = orc.vec(name, FR, x.reshape(1, 4), y.reshape(1, 4))[0]
        elif grp == "Pp":
            val[i] = orc.vec(name, FR, x, y)
        else:  # scalar x with polynomial y
            xs = bcast(x)
            if b == "Mul":
                val[i] = orc.vec(name, FR, xs, y)
            elif graph["poly_repr"] == "Coef":
                # only element 0 of the first chunk sees the scalar (kernel_gen.rs:503-534)
                out = y.copy() if b != "Sub" else orc.vec("orc_vec_sub", FR, zero, y)
                head = {"Add": orc.vec("orc_vec_add", FR, xs[:1], y[:1]),
                        "Sub": orc.vec("orc_vec_sub", FR, xs[:1], y[:1]),
                        "SubBy": orc.vec("orc_vec_sub", FR, y[:1], xs[:1])}[b]
                out[0] = head[0]
                val[i] = out
            else:
                val[i] = orc.vec(name, FR, y, xs) if b == "SubBy" else orc.vec(name, FR, xs, y)
        return val[i]

    return get


def kernel_meta(graph, name):
    inputs = graph["inputs"]
    V = graph["g"][0]
    inplace = {V[o]["op"]["Output"]["in_node"] for o in graph["outputs"] if V[o]["op"]["Output"]["in_node"] is not None}
    nv = sum(1 for i in inputs if i not in inplace)
    return fused.FusedKernelMeta(name, nv, len(graph["outputs"]), fused.LARGE_DIR)


# ---------------------------------------------------------------- CPU: the generator's plan


@pytest.mark.parametrize("seed,coef", CASES)
def test_large_graph_plan(seed, coef):
    gj = fused.random_graph(seed, fused.LARGE_VERTICES, coef=coef)
    g = json.loads(gj)
    assert len(g["g"][0]) >= 800
    src = fused.codegen(Libs(), gj, "large_plan")
    head = src.splitlines()[0]
    # a few hundred vertices reach the outputs, dozens of scalars go to the table, some values spill
    reached = int(head.split("(")[1].split(" vertices")[0])
    assert reached >= 500, head
    assert "in a table" in head and " 0 spill slots" not in head, head
    assert "_scalars_kernel" in src and "spill_store" in src and "spill_load" in src
    assert "amdgpu_waves_per_eu" in src


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("seed,coef", CASES)
def test_large_graph_matches_interpreter(libs, stream, seed, coef):
    name = fused.large_kernel_name(seed, coef)
    gj = fused.random_graph(seed, fused.LARGE_VERTICES, coef=coef)
    g = json.loads(gj)
    V = g["g"][0]
    meta = kernel_meta(g, name)
    k = fused.FusedKernel(libs, meta)
    assert k.local_bytes > 0
    n = 3000  # several waves, not a multiple of the workgroup
    polys, scalars, bufs = {}, {}, []
    var, var_ids = [], []
    outs = g["outputs"]
    inplace = {V[o]["op"]["Output"]["in_node"]: o for o in outs if V[o]["op"]["Output"]["in_node"] is not None}
    for j, i in enumerate(g["inputs"]):
        is_scalar = V[i]["op"]["Input"]["typ"] == "Scalar"
        if is_scalar:
            scalars[i] = random_elements(1, 10 * seed + j)[0]
        else:
            polys[i] = random_elements(n, 1000 * seed + j)
    get = interpret(g, polys, scalars, n)
    for i in g["inputs"]:
        if i in inplace:
            continue
        if i in scalars:
            b = hip.DeviceBuffer(32)
            b.upload(scalars[i])
            bufs.append(b)
            var.append(Scalar(b.addr))
        else:
            b, arr = device_array(n)
            arr.rotate(7 * (len(bufs) + 1))  # rotated operands, through the PolyPtr map
            cpu2gpu(polys[i], arr, b)
            bufs.append(b)
            var.append(arr)
    mut, out_arrays = [], []
    for o in outs:
        op = V[o]["op"]["Output"]
        if op["typ"] == "Scalar":
            b = hip.DeviceBuffer(32)
            bufs.append(b)
            mut.append(Scalar(b.addr))
            out_arrays.append(("s", b, None))
        else:
            b, arr = device_array(n)
            if op["in_node"] is not None:
                cpu2gpu(polys[op["in_node"]], arr, b)  # the in-place input, read through this slot
            bufs.append(b)
            mut.append(arr)
            out_arrays.append(("p", b, arr))
    arg = hip.DeviceBuffer(fused.arg_buffer_bytes(meta))
    local = hip.DeviceBuffer(k.local_bytes)
    print(f"large fused graph seed={seed} coef={coef} local={k.local_bytes}", flush=True)
    k.get_fn().f([GpuBuffer(arg.addr, arg.nbytes), GpuBuffer(local.addr, k.local_bytes)] + mut, [stream] + var)
    for o, (kind, b, arr) in zip(outs, out_arrays):
        want = get(V[o]["op"]["Output"]["store_node"])
        if kind == "s":
            got = b.download(np.uint64, 4)
            want = want if want.ndim == 1 else want[0]
            assert list(got) == list(want), o
        else:
            got = gpu2cpu(arr, b)
            want = want if want.ndim == 2 else np.broadcast_to(want, (n, 4))
            assert np.array_equal(got, want), o
