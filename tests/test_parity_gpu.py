"""FP64 GPU step against the CPU binary64 oracle on identical seeded inputs.

Bar: every order-parameter and composition field within 1e-10
max-abs after the stated number of steps; the displacement records, elastic
energy and intermediate fields of the first step to the same bar relative to
their scale.  MIXED is checked against the FP64 GPU path: relative L2
<= 1e-4 per field.
"""
import numpy as np
import pytest

from meumapps_ss_amd.params import MM_FP64, MM_MIXED
from meumapps_ss_amd.sim import Simulation
from oracle import Oracle
from parity_cases import CASES, make_case

pytestmark = pytest.mark.gpu

TOL = 1e-10
LAGGED = ("glob_min", "glob_max", "phi_max", "solid_tot")


def _rel(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


@pytest.mark.parametrize("name", list(CASES))
def test_first_step_intermediates(gpu_lib, name):
    p, phi, cons = make_case(name)
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(1)
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        rows = s.step(1)
        g = o.get_work(0)
        V = p.vartot
        gp = s.field("grad_phi")  # [v*3+d]
        gp_o = np.stack([g[d * V + v] for v in range(V) for d in range(3)])
        assert _rel(gp, gp_o) < TOL, "grad phi"
        assert _rel(s.field("k_del_sq"), o.get_work(1)) < TOL, "k_del_sq"
        assert _rel(s.field("el_grad"), o.get_work(4)) < TOL, "el_grad"
        assert _rel(s.field("f_diff"), o.get_work(3)) < TOL, "f_diff"
        st = s.download()
        ro = o.get_state()
        for k in ("phi", "cons"):
            assert np.abs(st[k] - ro[k]).max() < TOL, k
        for k in ("u", "gradu", "el_en"):
            assert _rel(st[k], ro[k]) < TOL, k
    r, q = rows[0], ref_rows[0]
    assert r["n_elastic"] == q["n_elastic"]
    for k in ("fint", "fch", "fw", "elen", "etot"):
        assert abs(r[k] - q[k]) <= 1e-9 * max(1.0, abs(q[k])), k
    assert abs(r["caviter"][0] - q["caviter"][0]) < 1e-13


@pytest.mark.parametrize("name,nsteps", [("v3_tetra", 40), ("v3_inhomog", 20), ("v2_twophase", 30),
                                         ("v3_noise", 60), ("v1_e2", 20), ("v12_rot", 10),
                                         ("v3_n24", 20), ("v1_tetC", 10), ("v1_hexC", 10), ("v1_trigC", 10)])
def test_multi_step(gpu_lib, name, nsteps):
    p, phi, cons = make_case(name, seed=1)
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(nsteps)
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        rows = s.step(nsteps)
        st = s.download()
    ro = o.get_state()
    for k in ("phi", "cons"):
        err = np.abs(st[k] - ro[k]).max()
        assert err < TOL, f"{k} max-abs {err:.3e}"
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]
    # the fort.98 rows: energies, and the lagged check-loop columns
    # (Multi_MEUMAPPS.f:905-906, from the check loop :1016-1062 of the step before)
    for r, q in zip(rows, ref_rows):
        for k in ("fint", "fch", "fw", "elen", "etot"):
            assert abs(r[k] - q[k]) <= 1e-9 * max(1.0, abs(q[k])), (k, q["step"])
        for k in LAGGED:
            assert abs(r[k] - q[k]) < TOL, (k, q["step"])
    # Cahn-Hilliard conserves mass to round-off
    c0 = cons.mean(axis=(1, 2, 3))
    for r in rows:
        for e in range(p.numel):
            assert abs(r["caviter"][e] - c0[e]) < 1e-13


@pytest.mark.parametrize("name", ["v3_tetra", "v2_twophase"])
def test_mixed_rows_match_oracle(gpu_lib, name):
    """MIXED fort.98 / fort.99 rows against the FP64 oracle over 20 steps: the
    lagged columns glob_min, glob_max, phi_max, solid_tot (formed from the FP64
    mean composition plus the FP32 deviation, and FP32 phi) within 1e-5, the
    energies within 1e-5 relative, the mean composition within 1e-9."""
    p, phi, cons = make_case(name, seed=1)
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(20)
    with Simulation(p, MM_MIXED) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        rows = s.step(20)
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]
    for r, q in zip(rows, ref_rows):
        for k in LAGGED:
            assert abs(r[k] - q[k]) < 1e-5, (k, q["step"], r[k], q[k])
        for k in ("fint", "fch", "fw", "elen", "etot"):
            assert abs(r[k] - q[k]) <= 1e-5 * max(1.0, abs(q[k])), (k, q["step"])
        for e in range(p.numel):
            assert abs(r["caviter"][e] - q["caviter"][e]) < 1e-9


def test_nv12_model_read_from_input_files(gpu_lib, tmp_path):
    """The nv = 12 model as a run sets it up: Input.txt and sym_var.inp written
    to disk, read back by read_input_txt (the rotations record by record, as
    InitTransGrad reads them, ModInitials.f:200-340), set up, and stepped on
    the GPU against the oracle."""
    from meumapps_ss_amd.params import read_input_txt, setup, write_input_txt
    from parity_cases import z_rotations

    from meumapps_ss_amd.configs import model_inputs

    p0, phi, cons = make_case("v12_rot", seed=2)
    inp = model_inputs(16, 12, (0.02, -0.01, 0.005), 0.005)
    write_input_txt(inp, str(tmp_path / "Input.txt"))
    with open(tmp_path / "sym_var.inp", "w") as f:
        for s_ in z_rotations():
            for row in s_:
                f.write(" ".join(repr(float(x)) for x in row) + "\n")
    p = setup(read_input_txt(str(tmp_path / "Input.txt")))
    assert np.array_equal(p.trans_ppt_array(), p0.trans_ppt_array())
    assert np.array_equal(np.ctypeslib.as_array(p.grad_coeff), np.ctypeslib.as_array(p0.grad_coeff))
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(4)
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        rows = s.step(4)
        st = s.download()
    ro = o.get_state()
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]
    for k in ("phi", "cons"):
        assert np.abs(st[k] - ro[k]).max() < TOL, k


def test_elastep0_later(gpu_lib):
    """Elasticity switched on after two steps (step < elastep0 path, then zeroth order)."""
    p, phi, cons = make_case("v3_tetra", seed=2)
    p.elastep0 = p.step + 2
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    o.step(5)
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        s.step(5)
        st = s.download()
    ro = o.get_state()
    for k in ("phi", "cons"):
        assert np.abs(st[k] - ro[k]).max() < TOL
    assert _rel(st["u"], ro["u"]) < TOL


def test_restart_warm_start(gpu_lib):
    """Upload u and its gradients (restart records) with step > elastep0."""
    p, phi, cons = make_case("v3_tetra", seed=3)
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    o.step(3)
    mid = o.get_state()
    o.step(3)
    ro = o.get_state()
    p2 = p
    p2.step = p.step + 3
    with Simulation(p2, MM_FP64) as s:
        s.upload(phi=mid["phi"], cons=mid["cons"], u=mid["u"], gradu=mid["gradu"], el_en=mid["el_en"])
        s.step(3)
        st = s.download()
    for k in ("phi", "cons"):
        assert np.abs(st[k] - ro[k]).max() < TOL
    assert _rel(st["gradu"], ro["gradu"]) < TOL


def test_mixed_vs_fp64(gpu_lib):
    p, phi, cons = make_case("v3_noise", seed=4)
    out = {}
    for prec in (MM_FP64, MM_MIXED):
        with Simulation(p, prec) as s:
            s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
            rows = s.step(40)
            out[prec] = (s.download(), rows)
    a, b = out[MM_FP64][0], out[MM_MIXED][0]
    for k in ("phi", "cons"):
        l2 = np.linalg.norm(b[k] - a[k]) / np.linalg.norm(a[k])
        assert l2 < 1e-4, f"{k} rel L2 {l2:.2e}"
    # FP64 k=0 handling keeps the mean composition to round-off
    c0 = cons.mean()
    for r in out[MM_MIXED][1]:
        assert abs(r["caviter"][0] - c0) < 1e-10


@pytest.mark.parametrize("name", ["v3_inhomog", "v2_nonsq", "v3_n24", "v3_n96", "v3_z192", "v3_z512", "v1_e2",
                                  "v12_rot", "v2_twophase"])
def test_mixed_cases(gpu_lib, name):
    """MIXED against FP64 on the other kernel combinations: inhomogeneous
    moduli (unfused x path), a non-cubic grid, radix-3 lengths (6 points per
    thread in the fused z kernels at 24 and 96, the 512-thread elastic shape
    at 192), binary32 z columns of 512 held by single wavefronts, three
    solutes, twelve variants (unfused x path), three phases."""
    p, phi, cons = make_case(name, seed=6)
    out = {}
    for prec in (MM_FP64, MM_MIXED):
        with Simulation(p, prec) as s:
            s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
            rows = s.step(10)
            out[prec] = (s.download(), rows)
    a, b = out[MM_FP64][0], out[MM_MIXED][0]
    for k in ("phi", "cons", "u"):
        if not np.any(a[k]):  # no eigenstrain: zero displacement in both
            assert not np.any(b[k]), k
            continue
        l2 = np.linalg.norm(b[k] - a[k]) / np.linalg.norm(a[k])
        assert l2 < 1e-4, f"{k} rel L2 {l2:.2e}"
    # the mean is held in FP64; what is left is the FP32 rounding of the
    # stored deviation, at most 2^-24 of its largest value
    c0 = cons.mean(axis=(1, 2, 3))
    for r in out[MM_MIXED][1]:
        for e in range(p.numel):
            bound = 2.0 ** -24 * np.abs(cons[e] - c0[e]).max()
            assert abs(r["caviter"][e] - c0[e]) < bound


def test_mixed_mass_does_not_drift(gpu_lib):
    """MIXED over 200 steps: the k = 0 mode of the stored deviation c - cbar
    is reset by the Cahn-Hilliard update, so the FP32 rounding of the
    transforms does not accumulate into the mean (left to accumulate it
    reaches ~4e-9 here by step 200; reset it stays below ~4e-10)."""
    p, phi, cons = make_case("v3_tetra", seed=6)
    c0 = cons.mean()
    with Simulation(p, MM_MIXED) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        rows = s.step(200)
    drift = np.array([r["caviter"][0] - c0 for r in rows])
    assert np.abs(drift).max() < 2.0 ** -26 * np.abs(cons - c0).max()


@pytest.mark.parametrize("name", ["v3_tetra", "v3_n24", "v3_inhomog"])
def test_strain_field(gpu_lib, name):
    """The strain field (kept as x half-spectra between steps on the fused x
    path) against the engineering strain of the oracle's displacement
    gradients after several steps."""
    p, phi, cons = make_case(name, seed=5)
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    o.step(4)
    g = o.get_state()["gradu"]  # g[3*d + i] = d_{d+1} u_{i+1}
    ref = np.stack([g[0], g[4], g[8], g[5] + g[7], g[2] + g[6], g[1] + g[3]])
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        s.step(4)
        eps = s.field("strain")
        # the real fields are formed on demand; stepping on from here is unaffected
        s.step(2)
        st = s.download()
    assert _rel(eps.reshape(ref.shape), ref) < TOL
    o.step(2)
    ro = o.get_state()
    for k in ("phi", "cons"):
        assert np.abs(st[k] - ro[k]).max() < TOL, k


def test_state_save_load_roundtrip(gpu_lib):
    """mm_state_save / mm_state_load on the fused x path (strain kept as x
    half-spectra): a reloaded state steps to the same fields within round-off."""
    import torch
    from meumapps_ss_amd import _lib as L
    import ctypes
    p, phi, cons = make_case("v3_tetra", seed=6)
    with Simulation(p, MM_FP64) as s:
        lib = s.lib
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        s.step(3)
        nb = ctypes.c_int64()
        L.check(lib.mm_state_bytes(s.ctx, ctypes.byref(nb)), "mm_state_bytes")
        buf = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
        L.check(lib.mm_state_save(s.ctx, ctypes.c_void_p(buf.data_ptr())), "mm_state_save")
        s.step(3)
        a = s.download()
        L.check(lib.mm_state_load(s.ctx, ctypes.c_void_p(buf.data_ptr())), "mm_state_load")
        s.step(3)
        b = s.download()
    for k in ("phi", "cons"):
        assert np.abs(a[k] - b[k]).max() < 1e-12, k
    assert _rel(a["u"], b["u"]) < 1e-12


def test_mixed_state_load_restores_the_mean_composition(gpu_lib):
    """MIXED stores c - cbar with cbar the FP64 mean of the upload.  A state
    saved in one context and loaded after an upload of other fields (another
    cbar), or into a second context, steps exactly like the uninterrupted
    run: the record carries cbar."""
    import ctypes

    import torch
    from meumapps_ss_amd import _lib as L
    p, phi, cons = make_case("v3_tetra", seed=7)
    zeros = dict(u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
    with Simulation(p, MM_MIXED) as s, Simulation(p, MM_MIXED) as t:
        lib = s.lib
        s.upload(phi=phi, cons=cons, **zeros)
        s.step(3)
        nb = ctypes.c_int64()
        L.check(lib.mm_state_bytes(s.ctx, ctypes.byref(nb)), "mm_state_bytes")
        buf = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
        L.check(lib.mm_state_save(s.ctx, ctypes.c_void_p(buf.data_ptr())), "mm_state_save")
        s.step(3)
        a = s.download()
        rows_a = s.step(1)
        # other fields, another mean composition, then the saved state
        s.upload(phi=0.5 * phi, cons=cons + 0.05, **zeros)
        L.check(lib.mm_state_load(s.ctx, ctypes.c_void_p(buf.data_ptr())), "mm_state_load")
        s.step(3)
        b = s.download()
        rows_b = s.step(1)
        # a second context that never saw the original upload
        t.upload(phi=phi, cons=cons + 0.1, **zeros)
        L.check(lib.mm_state_load(t.ctx, ctypes.c_void_p(buf.data_ptr())), "mm_state_load")
        t.step(3)
        c = t.download()
    for k in ("phi", "cons"):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k
    assert rows_a[0]["caviter"] == rows_b[0]["caviter"]


@pytest.mark.parametrize("form", ["nonsymmetric", "coupled"])
def test_homogeneous_moduli_forms_match_oracle(gpu_lib, form):
    """The fused x kernels read CEla as a symmetric matrix (checked bitwise at
    context creation) and, for the block form of the cubic, tetragonal and
    hexagonal classes, skip its zero terms.  Two homogeneous moduli outside
    that form, in every phase with the compliance kept as the inverse:
    nonsymmetric (C12 != C21) runs the unfused polarization and elener passes,
    coupled (symmetric, C14 = C41 != 0) the fused kernels on the full matrix.
    Both must match the oracle step for step."""
    p, phi, cons = make_case("v3_tetra", seed=3)
    for ph in range(p.numph):
        C = np.array([[p.cela[ph * 36 + j * 6 + i] for j in range(6)] for i in range(6)])
        if form == "nonsymmetric":
            C[0, 1] += 3.0
        else:
            C[0, 3] = C[3, 0] = 5.0
        S = np.linalg.inv(C)
        for i in range(6):
            for j in range(6):
                p.cela[ph * 36 + j * 6 + i] = C[i, j]
                p.sela[ph * 36 + j * 6 + i] = S[i, j]
    nsteps = 12
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(nsteps)
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
        rows = s.step(nsteps)
        st = s.download()
    ro = o.get_state()
    for k in ("phi", "cons", "u"):
        err = np.abs(st[k] - ro[k]).max()
        assert err < TOL, f"{k} max-abs {err:.3e}"
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]


def test_backtracking_matches_oracle(gpu_lib):
    """Soft precipitates (0.24 x the matrix moduli) make the first elastic loop
    diverge: the error grows, cfac is cut by newfac (backtracking,
    ModElastic.f:438-481) and ramped back; 271 iterations in the first step.
    The iteration counts and the fields match the oracle."""
    from meumapps_ss_amd.configs import model_inputs
    from meumapps_ss_amd.params import setup
    from parity_cases import smooth_field
    n, f = 16, 0.24
    inp = model_inputs(n, 3, (0.02, -0.01, 0.01), 0.005,
                       cela2={(1, 1): 250 * f, (1, 2): 150 * f, (4, 4): 100 * f})
    p = setup(inp)
    rng = np.random.default_rng(3)
    shape = (n, n, n)
    phi = np.stack([smooth_field(rng, shape, -0.15, 1.1) for _ in range(3)])
    cons = np.stack([smooth_field(rng, shape, 0.3, 0.6)])
    o = Oracle(p, 4)
    o.set_state(phi=phi, cons=cons)
    ref_rows = o.step(3)
    assert ref_rows[0]["n_elastic"] > 10  # the backtracking path ran
    with Simulation(p, MM_FP64) as s:
        s.upload(phi=phi, cons=cons, u=np.zeros((3, *shape)), gradu=np.zeros((9, *shape)))
        rows = s.step(3)
        st = s.download()
    assert [r["n_elastic"] for r in rows] == [q["n_elastic"] for q in ref_rows]
    assert [r["cfac"] for r in rows] == [q["cfac"] for q in ref_rows]
    ro = o.get_state()
    for k in ("phi", "cons"):
        assert np.abs(st[k] - ro[k]).max() < TOL, k


@pytest.mark.parametrize("name,prec", [("v3_tetra", MM_FP64), ("v3_n24", MM_MIXED), ("v1_e2", MM_FP64)])
def test_second_stream_gives_identical_bits(gpu_lib, name, prec, monkeypatch):
    """The Cahn-Hilliard update on its own stream (the default for eager steps)
    and on the main stream (MM_CH_STREAM=0) give the same bits."""
    p, phi, cons = make_case(name, seed=8)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("MM_CH_STREAM", flag)
        with Simulation(p, prec) as s:
            s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
            rows = s.step(8)
            out[flag] = (s.download(), [r["n_elastic"] for r in rows], [r["caviter"][0] for r in rows])
    a, b = out["0"], out["1"]
    for k in ("phi", "cons", "u", "gradu", "el_en"):
        assert np.array_equal(a[0][k], b[0][k]), k
    assert a[1] == b[1] and a[2] == b[2]


@pytest.mark.parametrize("prec", [MM_FP64, MM_MIXED])
def test_fixed_plans_are_bitwise_repeatable(gpu_lib, prec, monkeypatch):
    """Two contexts with the default FFT plans (MM_AUTOTUNE=0) give the same
    bits over 20 steps of the 64^3 one-variant workload (fixed-order
    reductions, no atomics; the Cahn-Hilliard update on its own stream)."""
    from meumapps_ss_amd.configs import initial_fields, workload

    monkeypatch.setenv("MM_AUTOTUNE", "0")
    w = workload(1, n=64)
    p = w.params()
    phi, cons = initial_fields(w)
    outs = []
    for _ in range(2):
        with Simulation(p, prec) as s:
            s.upload(phi=phi, cons=cons, u=np.zeros((3, *phi.shape[1:])), gradu=np.zeros((9, *phi.shape[1:])))
            rows = s.step(21)
            outs.append((s.download(), [(r["fint"], r["fch"], r["elen"], r["n_elastic"]) for r in rows]))
    for k in ("phi", "cons", "u", "gradu", "el_en"):
        assert np.array_equal(outs[0][0][k], outs[1][0][k]), k
    assert outs[0][1] == outs[1][1]
