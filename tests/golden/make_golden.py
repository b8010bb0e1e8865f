"""Writes the golden fixtures under tests/golden/fixtures/ by running the reference itself.

Every fixture is produced by oracle/_ref/ref_golden, i.e. by the reference's own CPU sources
(operators/, solver/, linsolve/) compiled from the reference checkout and driven through their
public C++ API. A fixture holds the inputs (seeded numpy data), the parameters, and the
reference's outputs; nothing of the reference's source travels with it.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ofref, refrun  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures")
CELL, FX, FY, FZ = ofref.CELL, ofref.FACE_X, ofref.FACE_Y, ofref.FACE_Z
P, NS, PV, SYM, FP = ofref.PERIODIC, ofref.NO_SLIP, ofref.PRESCRIBED, ofref.SYMMETRY, ofref.FIXED_PRESSURE

BC_SETS = {
    "periodic3d": dict(bc=[P] * 6),
    "tg2d": dict(bc=[P, P, P, P, SYM, SYM]),
    "cavity": dict(bc=[NS, NS, NS, PV, SYM, SYM], bcvel=[0] * 9 + [1.0, 0, 0] + [0] * 6),
    "channel": dict(bc=[FP, FP, NS, PV, SYM, SYM], bcvel=[0] * 9 + [0.7, 0, 0] + [0] * 6,
                    bcp=[0.8, 0.0, 0, 0, 0, 0]),
    "walls3d": dict(bc=[NS, NS, SYM, PV, NS, NS], bcvel=[0] * 9 + [0.5, 0.0, -0.25] + [0] * 6),
}


def shapes(g, locs):
    return [ofref.storage_shape(g["nx"], g["ny"], g["nz"], loc) for loc in locs]


VEL = [FX, FY, FZ]


def rnd(rng, g, locs, scale=1.0, zeros_frac=0.0):
    out = []
    for shp in shapes(g, locs):
        a = rng.standard_normal(shp) * scale
        if zeros_frac:
            a[rng.random(shp) < zeros_frac] = 0.0
        out.append(a)
    return out


def save(name, params, inputs: dict, outputs: dict, meta: dict):
    os.makedirs(OUT, exist_ok=True)
    arrays = {f"in_{k}": v for k, v in inputs.items()}
    arrays.update({f"out_{k}": v for k, v in outputs.items()})
    clean_meta = {}
    for k, v in meta.items():
        if isinstance(v, np.ndarray):
            arrays[f"meta_{k}"] = v
        else:
            clean_meta[k] = v
    arrays["params"] = np.array(json.dumps(params))
    arrays["meta"] = np.array(json.dumps(clean_meta))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
    print(f"{name}: {', '.join(sorted(arrays))}")


def grid(nx, ny, nz, dx, dy, dz):
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz)


def tg_params(nx, ny, nz, viscosity, cfl, steps, tol=1e-12, maxit=400):
    L = 2 * np.pi
    inv = nx / L + ny / L + (nz / L if nz > 1 else 0.0)
    upper = cfl / inv
    return dict(nx=nx, ny=ny, nz=nz, viscosity=viscosity, cfl=cfl, final_time=steps * upper * (1 - 1e-12),
                tol=tol, maxit=maxit)


def main():
    rng = np.random.default_rng(20261203)

    # -- operators: gradient, divergence, Laplacian (operators/discrete_operators.cpp)
    for name, g in [("operators_3d", grid(6, 5, 4, 0.3, 0.4, 0.5)), ("operators_2d", grid(8, 6, 1, 0.25, 0.3, 1.0))]:
        p = rnd(rng, g, [CELL])[0]
        vel = rnd(rng, g, VEL)
        outs, meta = refrun.run("operators", [p] + vel, shapes(g, VEL + [CELL, CELL] + VEL), **g)
        save(name, g, dict(p=p, u=vel[0], v=vel[1], w=vel[2]),
             dict(gx=outs[0], gy=outs[1], gz=outs[2], div=outs[3], lap_p=outs[4], lap_u=outs[5],
                  lap_v=outs[6], lap_w=outs[7]), meta)

    # -- advection (solver/momentum_terms.cpp), all three schemes, including zero/tiny velocities
    for scheme, sname in [(0, "tvd"), (1, "upwind"), (2, "central")]:
        for name, g in [("3d", grid(6, 5, 4, 0.3, 0.4, 0.5)), ("2d", grid(10, 8, 1, 0.2, 0.25, 1.0))]:
            vel = rnd(rng, g, VEL, zeros_frac=0.15)
            vel[0].flat[::7] = 1e-15
            vel[1].flat[::5] = vel[1].flat[1::5][: vel[1].flat[::5].size]
            params = dict(g, scheme=scheme)
            outs, meta = refrun.run("advection", vel, shapes(g, VEL), **params)
            save(f"advection_{name}_{sname}", params, dict(u=vel[0], v=vel[1], w=vel[2]),
                 dict(au=outs[0], av=outs[1], aw=outs[2]), meta)

    # -- diffusion + CFL
    g = grid(6, 5, 4, 0.3, 0.4, 0.5)
    vel = rnd(rng, g, VEL)
    outs, meta = refrun.run("diffusion", vel, shapes(g, VEL), nu=0.07, **g)
    save("diffusion_3d", dict(g, nu=0.07), dict(u=vel[0], v=vel[1], w=vel[2]),
         dict(du=outs[0], dv=outs[1], dw=outs[2]), meta)
    _, meta = refrun.run("cfl", vel, (), dt=0.013, **g)
    save("cfl_3d", dict(g, dt=0.013), dict(u=vel[0], v=vel[1], w=vel[2]), {}, meta)

    # -- velocity boundary conditions (solver/projection.cpp:207-462,966-971)
    for bname in ["periodic3d", "tg2d", "cavity", "channel", "walls3d"]:
        g = grid(6, 5, 1 if bname in ("tg2d", "cavity", "channel") else 4, 0.3, 0.4, 1.0 if bname in ("tg2d", "cavity", "channel") else 0.5)
        vel = rnd(rng, g, VEL)
        params = dict(g, **BC_SETS[bname])
        outs, meta = refrun.run("vbc", vel, shapes(g, VEL), **params)
        save(f"vbc_{bname}", params, dict(u=vel[0], v=vel[1], w=vel[2]),
             dict(u=outs[0], v=outs[1], w=outs[2]), meta)

    # -- pressure and total-pressure ghosts (solver/projection.cpp:302-382,973-1026)
    g = grid(6, 5, 4, 0.3, 0.4, 0.5)
    for pname, pb in [("periodic3d", dict(pbc=[2] * 6)),
                      ("mixed", dict(pbc=[0, 1, 2, 2, 1, 0], pbcval=[0, 0.3, 0, 0, -0.2, 0],
                                     pbcgrad=[0.5, 0, 0, 0, 0, -1.5]))]:
        p = rnd(rng, g, [CELL])[0]
        params = dict(g, **pb)
        outs, meta = refrun.run("pbc", [p], shapes(g, [CELL]), **params)
        save(f"pbc_{pname}", params, dict(p=p), dict(p=outs[0]), meta)
    for bname in ["cavity", "channel", "walls3d", "periodic3d"]:
        gg = g if bname in ("walls3d", "periodic3d") else grid(6, 5, 1, 0.3, 0.4, 1.0)
        src = rnd(rng, gg, VEL)
        p = rnd(rng, gg, [CELL])[0]
        params = dict(gg, **BC_SETS[bname])
        outs, meta = refrun.run("tpbc", src + [p], shapes(gg, [CELL]), **params)
        save(f"tpbc_{bname}", params, dict(su=src[0], sv=src[1], sw=src[2], p=p), dict(p=outs[0]), meta)

    # -- cavity predictor RHS (solver/lid_driven_cavity.cpp:489-518)
    g = grid(8, 6, 1, 1 / 8, 1 / 6, 1.0)
    for has_prev in (0, 1):
        cur = rnd(rng, g, VEL)
        p = rnd(rng, g, [CELL])[0]
        prev = rnd(rng, g, VEL)
        params = dict(g, nu=0.01, dt=0.004, has_prev=has_prev)
        outs, meta = refrun.run("cavity_rhs", cur + [p] + prev, shapes(g, VEL + VEL), **params)
        save(f"cavity_rhs_{'ab2' if has_prev else 'first'}", params,
             dict(u=cur[0], v=cur[1], w=cur[2], p=p, pu=prev[0], pv=prev[1], pw=prev[2]),
             dict(au=outs[0], av=outs[1], aw=outs[2], ru=outs[3], rv=outs[4], rw=outs[5]), meta)

    # -- ADI Helmholtz predictor (solver/projection.cpp:522-888,1028-1047)
    for bname, g, alpha in [("periodic3d", grid(6, 5, 4, 0.3, 0.4, 0.5), 0.011),
                            ("tg2d", grid(8, 6, 1, 0.25, 0.3, 1.0), 0.02),
                            ("cavity", grid(8, 6, 1, 1 / 8, 1 / 6, 1.0), 0.0375),
                            ("channel", grid(8, 6, 1, 1 / 8, 1 / 6, 1.0), 0.01),
                            ("walls3d", grid(6, 5, 4, 0.3, 0.4, 0.5), 0.02)]:
        rhs = rnd(rng, g, VEL)
        params = dict(g, alpha=alpha, **BC_SETS[bname])
        outs, meta = refrun.run("adi", rhs, shapes(g, VEL), **params)
        save(f"adi_{bname}", params, dict(u=rhs[0], v=rhs[1], w=rhs[2]),
             dict(u=outs[0], v=outs[1], w=outs[2]), meta)

    # -- pressure RHS, velocity correction (solver/projection.cpp:1049-1095)
    for bname, g in [("periodic3d", grid(6, 5, 4, 0.3, 0.4, 0.5)), ("cavity", grid(8, 6, 1, 1 / 8, 1 / 6, 1.0))]:
        vel = rnd(rng, g, VEL)
        phi = rnd(rng, g, [CELL])[0]
        params = dict(g, dt=0.02, density=1.3, **BC_SETS[bname])
        outs, meta = refrun.run("bprhs", vel, shapes(g, [CELL]), **params)
        save(f"bprhs_{bname}", params, dict(u=vel[0], v=vel[1], w=vel[2]), dict(rhs=outs[0]), meta)
        outs, meta = refrun.run("correct", vel + [phi], shapes(g, VEL), **params)
        save(f"correct_{bname}", params, dict(u=vel[0], v=vel[1], w=vel[2], phi=phi),
             dict(u=outs[0], v=outs[1], w=outs[2]), meta)

    # -- projection (solver/projection.cpp:1097-1131) including the MGPCG solve
    for bname, g, tol in [("periodic3d", grid(8, 8, 8, 2 * np.pi / 8, 2 * np.pi / 8, 2 * np.pi / 8), 1e-12),
                          ("cavity", grid(16, 12, 1, 1 / 16, 1 / 12, 1.0), 1e-12),
                          ("channel", grid(16, 8, 1, 1 / 16, 1 / 8, 1.0), 1e-11)]:
        vel = rnd(rng, g, VEL)
        phi = np.zeros(shapes(g, [CELL])[0])
        params = dict(g, dt=0.01, density=1.0, tol=tol, maxit=400, **BC_SETS[bname])
        outs, meta = refrun.run("project", vel + [phi], shapes(g, [CELL] + VEL + [CELL]), **params)
        save(f"project_{bname}", params, dict(u=vel[0], v=vel[1], w=vel[2], phi=phi),
             dict(phi=outs[0], u=outs[1], v=outs[2], w=outs[3], prhs=outs[4]), meta)

    # -- MGPCG Poisson (linsolve/poisson_solver.cpp:678-809), all BC families
    for pname, g, pb, tol in [
        ("periodic3d", grid(8, 8, 8, 0.5, 0.5, 0.5), dict(pbc=[2] * 6), 1e-12),
        ("periodic2d", grid(16, 16, 1, 0.4, 0.4, 1.0), dict(pbc=[2, 2, 2, 2, 0, 0]), 1e-12),
        ("dirichlet2d", grid(16, 16, 1, 1 / 16, 1 / 16, 1.0),
         dict(pbc=[1, 1, 1, 1, 0, 0], pbcval=[0.1, -0.2, 0.3, 0.0, 0, 0]), 1e-12),
        ("neumann2d", grid(12, 12, 1, 1 / 12, 1 / 12, 1.0),
         dict(pbc=[0] * 6, pbcgrad=[0.2, -0.1, 0.0, 0.3, 0, 0]), 1e-12),
        ("mixed3d", grid(8, 8, 4, 0.25, 0.2, 0.3), dict(pbc=[2, 2, 1, 1, 0, 0], pbcval=[0, 0, 0.5, 0, 0, 0]), 1e-12),
        ("odd3d", grid(12, 6, 5, 0.25, 0.2, 0.3), dict(pbc=[2, 2, 2, 2, 0, 0]), 1e-12),
    ]:
        rhs = rnd(rng, g, [CELL])[0]
        p0 = rnd(rng, g, [CELL], scale=0.1)[0] if pname == "mixed3d" else np.zeros(shapes(g, [CELL])[0])
        params = dict(g, tol=tol, maxit=400, **pb)
        outs, meta = refrun.run("poisson", [rhs, p0], shapes(g, [CELL]), **params)
        save(f"poisson_{pname}", params, dict(rhs=rhs, p=p0), dict(p=outs[0]), meta)
        if pname in ("dirichlet2d", "neumann2d"):
            outs, meta = refrun.run("poisson_rhs", [p0 + rhs], shapes(g, [CELL]), **params)
            save(f"poisson_rhs_{pname}", params, dict(p=p0 + rhs), dict(rhs=outs[0]), meta)

    # -- Taylor-Green driver (solver/taylor_green.cpp): init + step loop
    for name, tp in [("tg2d_16", tg_params(16, 16, 1, 0.01, 0.5, 3)),
                     ("tg3d_8", tg_params(8, 8, 8, 0.05, 0.5, 3)),
                     ("tg3d_12x10x8", tg_params(12, 10, 8, 0.02, 0.4, 2))]:
        g = dict(nx=tp["nx"], ny=tp["ny"], nz=tp["nz"])
        shp = shapes(g, VEL + [CELL] + VEL)
        outs, meta = refrun.run("tg_init", (), shp, **tp)
        save(f"{name}_init", tp, {}, dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3]), meta)
        steps = 3 if name != "tg3d_12x10x8" else 2
        outs, meta = refrun.run("tg_steps", (), shp, steps=steps, **tp)
        save(f"{name}_steps", dict(tp, steps=steps), {},
             dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3], au=outs[4], av=outs[5], aw=outs[6]), meta)

    # shifted (u = sin x cos y cos z) initial state injected through TaylorGreenState
    tp = tg_params(8, 8, 8, 1 / 1600, 0.5, 2)
    c = ofref.tg_config(8, 8, 8, tp["viscosity"], tp["cfl"], tp["final_time"], tp["maxit"], tp["tol"])
    s0 = ofref.tg_init(c, convention=1)
    g = dict(nx=8, ny=8, nz=8)
    shp = shapes(g, VEL + [CELL] + VEL)
    ins = s0.vel + [s0.p] + s0.adv
    outs, meta = refrun.run("tg_steps", ins, shp, steps=2, has_prev=0, **tp)
    save("tg3d_8_shifted_steps", dict(tp, steps=2, has_prev=0),
         dict(u=ins[0], v=ins[1], w=ins[2], p=ins[3], au=ins[4], av=ins[5], aw=ins[6]),
         dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3], au=outs[4], av=outs[5], aw=outs[6]), meta)

    # non-converged projection must be rejected with the reference's message
    tp = tg_params(16, 16, 1, 0.01, 0.5, 1, tol=1e-30, maxit=1)
    _, meta = refrun.run("tg_steps", (), (), steps=1, **tp)
    save("tg_nonconverged", dict(tp, steps=1), {}, {}, meta)

    # -- lid-driven cavity driver (solver/lid_driven_cavity.cpp)
    cp = dict(nx=16, ny=16, reynolds=100.0, lid=1.0, cfl=0.5, tol=1e-12, maxit=400)
    g = dict(nx=16, ny=16, nz=1)
    shp = shapes(g, VEL + [CELL] + VEL)
    outs, meta = refrun.run("cavity_init", (), shp, **cp)
    save("cavity16_init", cp, {}, dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3]), meta)
    outs, meta = refrun.run("cavity_steps", (), shp, steps=6, **cp)
    save("cavity16_steps", dict(cp, steps=6), {},
         dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3], au=outs[4], av=outs[5], aw=outs[6]), meta)

    # -- manufactured-solution operator errors (solver/operator_verification.cpp)
    for n in (16, 32, 64):
        _, meta = refrun.run("mms", (), (), n=n)
        save(f"mms_{n}", dict(n=n), {}, {}, meta)

    channel_fixtures()
    io_fixtures()
    validation_fixture()


def channel_fixtures():
    """run_channel_flow (solver/channel_flow.cpp:432-541) on the reference's smoke and benchmark
    configurations (benchmarks/channel_*_smoke.cfg, channel_*_128.cfg): final metrics, the
    x-averaged streamwise profile and the validation figures; plus the non-converged rejection."""
    base = dict(viscosity=0.1, top=1.0, drop=0.8, cfl=0.5, maxit=200, tol=1e-10)
    for case in ("couette", "poiseuille"):
        for n, steps in ((32, 4), (128, 8)):
            cp = dict(base, case=case, nx=n, ny=n, steps=steps)
            outs, meta = refrun.run("channel", (), [(n,)], **cp)
            save(f"channel_{case}_{n}", cp, {}, dict(profile=outs[0]), meta)
    cp = dict(base, case="poiseuille", nx=32, ny=32, steps=4, maxit=1, tol=1e-30)
    _, meta = refrun.run("channel", (), (), **cp)
    save("channel_nonconverged", cp, {}, {}, meta)



def io_fixtures():
    """A cavity state after 6 steps with the SLVCHKP1 checkpoint and the legacy-VTK file the
    reference's own io/checkpoint.cpp and io/vtk_export.cpp write for it (file bytes kept as
    uint8 arrays)."""
    import tempfile

    cp = dict(nx=16, ny=16, reynolds=100.0, lid=1.0, cfl=0.5, tol=1e-10, maxit=200, max_steps=12, min_steps=1,
              steady_tolerance=1e-12, validate=0)
    g = dict(nx=16, ny=16, nz=1)
    shp = shapes(g, VEL + [CELL] + VEL)
    with tempfile.TemporaryDirectory() as td:
        ck, vtk = os.path.join(td, "state.chk"), os.path.join(td, "state.vtk")
        outs, meta = refrun.run("cavity_steps", (), shp, steps=6, ckpt=ck, vtk=vtk, **cp)
        ck_bytes = np.fromfile(ck, dtype=np.uint8)
        vtk_bytes = np.fromfile(vtk, dtype=np.uint8)
    save("io_cavity16", dict(cp, steps=6), {},
         dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3], au=outs[4], av=outs[5], aw=outs[6], checkpoint=ck_bytes,
              vtk=vtk_bytes), meta)



def validation_fixture():
    """The reference's published validation-suite results (validation/latest/*.csv, data files
    of the reference checkout): operator MMS errors, Taylor-Green temporal refinement and the
    benchmark-suite metrics."""
    import csv

    latest = os.path.join(os.environ.get("REF", "/root/reference"), "validation", "latest")
    tables = {}
    for name in ("operator_spatial", "taylor_green_temporal", "benchmark_summary"):
        with open(os.path.join(latest, name + ".csv"), newline="") as f:
            tables[name] = list(csv.DictReader(f))
    save("validation_latest", {"source": "validation/latest"}, {}, {}, {"tables": tables})


def long_run_fixtures():
    """Reference runs the small fixtures above cannot stand for:
    * tg3d_edges: two steps from a state whose edge and corner ghosts are nonzero, the values
      the 27-point factorisation correction reads (solver/taylor_green.cpp:145-225) and a BC
      fill never writes;
    * tg2d_128_c1: the 2D Taylor-Green benchmark configuration at full length (128^2,
      viscosity 0.01, CFL 0.5, 200 steps, MGPCG tolerance 1e-10) from the shifted initial
      field u = sin x cos y, v = -cos x sin y, injected through TaylorGreenState;
    * cavity_c2_128: the lid-driven cavity (128^2, Re 100) run to steady state by the rule of
      run_lid_driven_cavity at max|du| <= 1e-8 with Poisson tolerance 1e-10, and at the
      reference's own benchmark setting (1e-7, 1e-9): final metrics, centreline validation and
      fields."""
    rng = np.random.default_rng(20261204)
    tp = tg_params(10, 8, 6, 0.02, 0.4, 2)
    g = dict(nx=10, ny=8, nz=6)
    shp = shapes(g, VEL + [CELL] + VEL)
    ins = [0.5 * rng.standard_normal(s) for s in shp]
    outs, meta = refrun.run("tg_steps", ins, shp, steps=2, has_prev=1, **tp)
    save("tg3d_edges_steps", dict(tp, steps=2, has_prev=1),
         dict(u=ins[0], v=ins[1], w=ins[2], p=ins[3], au=ins[4], av=ins[5], aw=ins[6]),
         dict(u=outs[0], v=outs[1], w=outs[2], p=outs[3], au=outs[4], av=outs[5], aw=outs[6]), meta)

    tp = tg_params(128, 128, 1, 0.01, 0.5, 200, tol=1e-10, maxit=200)
    c = ofref.tg_config(128, 128, 1, tp["viscosity"], tp["cfl"], tp["final_time"], tp["maxit"], tp["tol"])
    s0 = ofref.tg_init(c, convention=1)
    g = dict(nx=128, ny=128, nz=1)
    shp = shapes(g, VEL + [CELL] + VEL)
    ins = s0.vel + [s0.p] + s0.adv
    outs, meta = refrun.run("tg_steps", ins, shp, steps=200, has_prev=0, timeout=1800, **tp)
    save("tg2d_128_c1", dict(tp, steps=200, has_prev=0), dict(u=ins[0], v=ins[1], p=ins[3]),
         dict(u=outs[0], v=outs[1], p=outs[3]), meta)

    cavity_fixtures()


def cavity_fixtures():
    import concurrent.futures as cf

    g = dict(nx=128, ny=128, nz=1)
    shp = shapes(g, VEL + [CELL] + VEL)
    cases = [("cavity_c2_128", 1e-8, 1e-10), ("cavity_bench_128", 1e-7, 1e-9)]

    def one(case):
        name, steady, tol = case
        cp = dict(nx=128, ny=128, reynolds=100.0, lid=1.0, cfl=0.5, max_steps=20000, min_steps=400,
                  steady_tolerance=steady, tol=tol, maxit=200, validate=1)
        outs, meta = refrun.run("cavity_run", (), shp, timeout=3600, **cp)
        return name, cp, outs, meta

    with cf.ThreadPoolExecutor(len(cases)) as ex:  # independent reference runs of several minutes
        for name, cp, outs, meta in ex.map(one, cases):
            save(name, cp, {}, dict(u=outs[0], v=outs[1], p=outs[3]), meta)


if __name__ == "__main__":
    if sys.argv[1:] == ["long"]:
        long_run_fixtures()
    elif sys.argv[1:] == ["cavity"]:
        cavity_fixtures()
    else:
        main()
        long_run_fixtures()
