"""Writes the golden fixtures in this folder from the reference's own pressure path, compiled from its sources
(``make -C oracle ref``; needs /root/reference, so it runs in the build container only).

Each fixture is a .npz (no pickles) holding the case description (dimension, GRID_SIZE, domain, refinement
intervals, cell types, right-hand side) and what the reference produced for it: the spacing tables
(MAC::BuildAxis/AddRing/GapTable), the unknown numbering, the CSR matrix (PressureSolver{2D,3D}::AssembleSystem,
CommitPressureMatrix), the singular flag, and solutions with iteration counts from SolvePressureSystem with
eigen_solver=CG/DIAGONAL and eigen_solver=AMG (AMGCL).

The ``fields_*.npz`` fixtures add seeded velocity and pressure fields (MAC layout) for some of those cases and
what the reference computes from them: the right-hand side with MAC::GetDivergencyAt and DirichletSum
(PressureSolver{2D,3D}::SolveScheme's expression), IncrementReferenceNorm with a previous pressure, and the
fields after ProjectPressure; then p after MAC::EnforcePressureBC with a known boundary pressure, after
SolveScheme's tail on a given vector standing for the solution, and (incremental) after
AccumulatePressureIncrement.

The ``floor_*.npz`` fixtures hold solves with a reference norm (the absolute floor of an incremental solve,
SolveEigenPressure) on some of those systems: their iteration counts and solutions.

    python tests/golden/make_fixtures.py            # everything
    python tests/golden/make_fixtures.py --floor    # the floor fixtures only
    python tests/golden/make_fixtures.py --level    # the level_*.npz fixtures only: LevelPressure with a scene's
                                                    # inlet levels and BuildWarmStartGuess
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import refcase as R  # noqa: E402

FLUID, SOLID, EMPTY, INFLOW, PRESSURE, OUTFLOW = 0, 1, 2, 3, 4, 5


def ring_mask(shape, dim):
    r = np.zeros(shape, bool)
    r[0] = r[-1] = True
    r[:, 0] = r[:, -1] = True
    if dim == 3:
        r[:, :, 0] = r[:, :, -1] = True
    return r


def layout(kind, shape, dim, rng):
    t = np.zeros(shape, np.uint8)
    ring = ring_mask(shape, dim)
    t[ring] = SOLID
    if kind == "neumann":
        pass
    elif kind == "outlet":
        t[:, -1][ring[:, -1]] = EMPTY
    elif kind == "pressure":
        t[ring] = PRESSURE
    elif kind == "mixed":
        codes = np.array([SOLID, EMPTY, PRESSURE, INFLOW, OUTFLOW, FLUID], np.uint8)
        t[ring] = codes[rng.integers(0, len(codes), int(ring.sum()))]
        r = rng.random(shape)
        inner = ~ring
        t[inner & (r < 0.1)] = SOLID
        t[inner & (r >= 0.1) & (r < 0.15)] = EMPTY
        t[inner & (r >= 0.15) & (r < 0.17)] = PRESSURE
    elif kind == "freesurface":
        inner = ~ring
        t[inner] = EMPTY
        Ny, Nx = shape[0], shape[1]
        if dim == 3:
            t[1:max(2, Ny // 2), 1:max(2, (2 * Nx) // 3), 1:-1] = FLUID
        else:
            t[1:max(2, Ny // 2), 1:max(2, (2 * Nx) // 3), 0] = FLUID
        idx = np.argwhere(inner)
        for c in idx[rng.integers(0, len(idx), 6)]:
            t[tuple(c)] = FLUID
    else:
        raise ValueError(kind)
    return t


H, RA, SY, RB = O.REFINE_HARD, O.REFINE_RAMP, O.REFINE_SYMMETRIC, O.REFINE_RAMP_BACK
REF = O.Refinement

CASES = [
    # name, dim, GRID_SIZE, domain, non_uniform, refinements (x, y, z), layout, compat
    ("u3_neumann_8", 3, 8, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "neumann", True),
    ("u3_outlet_9", 3, 9, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "outlet", False),
    ("u3_pressure_7", 3, 7, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "pressure", False),
    ("u3_mixed_10", 3, 10, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "mixed", False),
    ("u3_freesurface_10", 3, 10, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "freesurface", False),
    ("u3_box_6x8x10", 3, 6, (0, 1, 0, 8 / 6, 0, 10 / 6), False, ((), (), ()), "outlet", False),
    ("s3_hard_mixed", 3, 8, (0, 1, 0, 1, 0, 1), True,
     ((REF(0.25, 0.5, 2, H),), (REF(0.5, 1.0, 3, H),), ()), "mixed", False),
    ("s3_graded_outlet", 3, 10, (0, 1, 0, 1, 0, 1), True,
     ((REF(0.3, 0.7, 3, SY),), (REF(0.0, 0.4, 2, RA), REF(0.6, 1.0, 2, RB)), (REF(0.2, 0.5, 2, H), REF(0.5, 0.8, 3, RA))),
     "outlet", False),
    ("s3_equal_neumann", 3, 8, (0, 1, 0, 1, 0, 1), True, ((), (), ()), "neumann", True),
    ("u2_neumann_16", 2, 16, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "neumann", True),
    ("u2_pressure_12", 2, 12, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "pressure", False),
    ("u2_mixed_20", 2, 20, (0, 1, 0, 1, 0, 1), False, ((), (), ()), "mixed", False),
    ("s2_ramp_outlet", 2, 16, (0, 1, 0, 1, 0, 1), True, ((REF(0.1, 0.6, 3, RA),), (REF(0.0, 0.5, 2, SY),), ()),
     "outlet", False),
    ("s2_units_hard", 2, 20, (0, 1, 0, 1, 0, 1), True, ((REF(0.4, 0.6, 2, H),), (), ()), "freesurface", False),
]


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/ref_harness is not built: make -C oracle ref")
    for q, (name, dim, gs, dom, nu, refs, kind, compat) in enumerate(CASES):
        probe = R.run(R.RefCase(dim=dim, grid_size=gs, domain=dom, non_uniform=nu, refine=refs))
        shape = (probe["Ny"], probe["Nx"], probe["Nz"])
        rng = np.random.default_rng(1000 + q)
        types = layout(kind, shape, dim, rng)
        inner = types[1:-1, 1:-1, 1:-1] if dim == 3 else types[1:-1, 1:-1, :]
        n = int((inner == FLUID).sum())
        b = rng.standard_normal(n)
        if compat:
            b = O.compatibility(b)
        solves = [R.SolveSpec(R.CG, R.DIAGONAL, 1e-12, 5000), R.SolveSpec(R.AMG, R.DIAGONAL, 1e-10, 500)]
        case = R.RefCase(dim=dim, grid_size=gs, domain=dom, non_uniform=nu, refine=refs, types=types, compat=compat,
                         solves=solves, rhs=b)
        out = R.run(case, threads=1)
        flat_refs = []
        for a, axis in enumerate(refs):
            for r in axis:
                flat_refs.append([a, r.start, r.end, r.factor, r.mode])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            dim=dim, grid_size=gs, domain=np.array(dom, float), non_uniform=int(nu),
            refinements=np.array(flat_refs, float).reshape(-1, 5), compat=int(compat), types=types, rhs=b,
            dh=out["dh"], wx=out["wx"], wy=out["wy"], wz=out["wz"], gapx=out["gapx"], gapy=out["gapy"],
            gapz=out.get("gapz", np.zeros(0)), ref_non_uniform=int(out["non_uniform"]),
            n=out["n"], singular=int(out["singular"]), order=out["order"], rowptr=out["rowptr"], col=out["col"],
            val=out["val"],
            cg_x=out["solves"][0]["x"], cg_iters=out["solves"][0]["iters"], cg_tol=1e-12,
            amg_x=out["solves"][1]["x"], amg_iters=out["solves"][1]["iters"], amg_tol=1e-10,
        )
        print(f"{name}: dims {shape} n={out['n']} nnz={out['nnz']} singular={out['singular']} "
              f"cg_iters={out['solves'][0]['iters']} amg_iters={out['solves'][1]['iters']}")


FIELD_CASES = ["u3_mixed_10", "u3_pressure_7", "s3_hard_mixed", "s3_graded_outlet", "u2_mixed_20",
               "u2_pressure_12", "s2_ramp_outlet", "u3_neumann_8", "s3_equal_neumann", "u2_neumann_16",
               "u3_freesurface_10"]


def main_fields():
    for q, name in enumerate(FIELD_CASES):
        (_, dim, gs, dom, nu, refs, kind, compat) = next(c for c in CASES if c[0] == name)
        base = np.load(os.path.join(HERE, name + ".npz"))
        types = base["types"]
        grid = O.Grid(dim, types, base["wx"], base["wy"], base["wz"] if dim == 3 else np.array([float(base["dh"])]),
                      float(base["dh"]), bool(nu))
        nu_, nv_, nw_, np_ = O.field_sizes(grid)
        rng = np.random.default_rng(2000 + q)
        f = {"rho": float(rng.uniform(0.5, 2.0)), "dt": float(rng.uniform(1e-4, 1e-2)),
             "u": rng.standard_normal(nu_), "v": rng.standard_normal(nv_),
             "w": rng.standard_normal(nw_) if dim == 3 else None, "p": rng.standard_normal(np_),
             "p_old": rng.standard_normal(np_) if q % 2 == 0 else None, "t": float(rng.uniform(0.0, 2.0))}
        f["x_tail"] = rng.standard_normal(int(O.assemble(grid).n)) + 0.7
        out = R.run(R.RefCase(dim=dim, grid_size=gs, domain=dom, non_uniform=nu, refine=refs, types=types,
                              compat=True, fields=f), threads=1)
        np.savez_compressed(
            os.path.join(HERE, "fields_" + name + ".npz"),
            case=name, rho=f["rho"], dt=f["dt"], u=f["u"], v=f["v"], w=f["w"] if dim == 3 else np.zeros(0),
            p=f["p"], p_old=f["p_old"] if f["p_old"] is not None else np.zeros(0),
            rhs=out["rhs_built"], ref_norm=out["ref_norm"], u_proj=out["u"], v_proj=out["v"],
            w_proj=out["w"] if dim == 3 else np.zeros(0), t=f["t"], x_tail=f["x_tail"], p_bc=out["p_bc"],
            p_tail=out["p_tail"], p_acc=out["p_acc"] if f["p_old"] is not None else np.zeros(0))
        print(f"fields_{name}: n={out['n']} incremental={f['p_old'] is not None} ref_norm={out['ref_norm']:.6g}")


# SolveEigenPressure's absolute floor (PressureSettings.h:411-427) on uniform, stretched, singular and 2D systems:
# the early return (||b|| <= tol * ref: the guess, or 0, with 0 iterations), the min(0.5, tol * ref / ||b||) branch
# and its 0.5 cap, from a zero start and from a guess; eigen_solver = CG with the DIAGONAL preconditioner.
FLOOR_CASES = ["u3_outlet_9", "s3_graded_outlet", "u3_neumann_8", "u2_pressure_12", "u3_mixed_10"]


def floor_specs(b, guess, tol):
    bn = float(np.linalg.norm(b))
    return [
        ("early_zero", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=2.0 * bn / tol)),
        ("early_guess", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=1.25 * bn / tol, guess=guess)),
        ("min_zero", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=1e4 * bn)),
        ("min_guess", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=3e3 * bn, guess=guess)),
        ("cap_zero", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=0.999 * bn / tol)),
        ("cap_guess", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=0.99 * bn / tol, guess=guess)),
        ("plain_guess", R.SolveSpec(R.CG, R.DIAGONAL, tol, 5000, ref_norm=0.0, guess=guess)),
    ]


def main_floor():
    tol = 1e-10
    for q, name in enumerate(FLOOR_CASES):
        (_, dim, gs, dom, nu, refs, kind, compat) = next(c for c in CASES if c[0] == name)
        base = np.load(os.path.join(HERE, name + ".npz"))
        # a smooth right-hand side (low modes plus a little noise), so that one CG step leaves more than half of
        # it and the 0.5 cap decides when the floor's tolerance stops
        order = base["order"].astype(np.int64)
        Ny, Nx = base["types"].shape[:2]
        Nz = base["types"].shape[2]
        i, j, k = order // (Nx * Nz), (order // Nz) % Nx, order % Nz
        rng = np.random.default_rng(3000 + q)
        b = (np.cos(np.pi * (j + 0.5) / Nx) * np.cos(np.pi * (i + 0.5) / Ny) * (1.0 + 0.5 * np.cos(np.pi * (k + 0.5) / Nz))
             + 0.05 * rng.standard_normal(order.size))
        if compat:
            b = O.compatibility(b)
        grid = O.Grid(dim, base["types"], base["wx"], base["wy"],
                      base["wz"] if dim == 3 else np.array([float(base["dh"])]), float(base["dh"]), bool(nu))
        xs, _, _ = O.assemble(grid).cg(b, 1e-13, 100000)
        guess = xs + 0.2 * rng.standard_normal(b.size) * float(np.abs(xs).max())
        specs = floor_specs(b, guess, tol)
        out = R.run(R.RefCase(dim=dim, grid_size=gs, domain=dom, non_uniform=nu, refine=refs, types=base["types"],
                              compat=compat, solves=[s for _, s in specs], rhs=b), threads=1)
        arrays = {"case": name, "tol": tol, "rhs": b, "guess": guess, "names": np.array([k for k, _ in specs]),
                  "ref_norms": np.array([s.ref_norm for _, s in specs]),
                  "has_guess": np.array([s.guess is not None for _, s in specs], np.int32),
                  "iters": np.array([o["iters"] for o in out["solves"]], np.int32),
                  "x": np.stack([o["x"] for o in out["solves"]])}
        np.savez_compressed(os.path.join(HERE, "floor_" + name + ".npz"), **arrays)
        print(f"floor_{name}: n={out['n']} iters={arrays['iters'].tolist()}")


# LevelPressure's inflow shift (PressureSolver3D.cpp:15-30, :181-211) and BuildWarmStartGuess (:215-244): inlet
# faces with scene p= levels on an INFLOW ring (singular, compatibility on), 2D/3D, uniform/stretched, interior
# solids beside the inlets; a previous pressure for the warm start (zero in one case: no guess), and a non-singular
# case where LevelPressure does nothing and the guess keeps its mean.
SCENE_A = ('boundary xmin inlet u="1" p="0.5+y*z+t"\nboundary xmax inlet u="1" p="-0.25+x*y"\n'
           'boundary ymin noslip\nboundary ymax noslip\nboundary zmin inlet w="0.1" p="2*x-t"\nboundary zmax noslip\n')
SCENE_B = ('boundary xmin noslip\nboundary xmax noslip\nboundary ymin inlet v="1" p="1.5-x"\n'
           'boundary ymax inlet v="-1" p="0.75+x*z"\nboundary zmin noslip\nboundary zmax inlet w="-0.5" p="3+t*y"\n')
SCENE_2D = ('boundary xmin inlet u="1" p="0.5+y+t"\nboundary xmax noslip\nboundary ymin noslip\n'
            'boundary ymax inlet v="-1" p="2-x*x"\n')
LEVEL_CASES = [
    # name, dim, GRID_SIZE, non_uniform, refinements, scene, ring kind, interior solids, warm start
    ("level_u3_inflow", 3, 8, False, ((), (), ()), SCENE_A, "inflow", False, "random"),
    ("level_s3_inflow", 3, 10, True,
     ((REF(0.3, 0.7, 3, SY),), (REF(0.0, 0.4, 2, RA), REF(0.6, 1.0, 2, RB)), (REF(0.2, 0.5, 2, H),)), SCENE_B, "inflow",
     True, "random"),
    ("level_u3_solids", 3, 9, False, ((), (), ()), SCENE_A, "inflow", True, "zero"),
    ("level_u2_inflow", 2, 12, False, ((), (), ()), SCENE_2D, "inflow", True, "random"),
    ("level_u3_outlet", 3, 8, False, ((), (), ()), SCENE_A, "outlet", False, "random"),
]


def main_level():
    for q, (name, dim, gs, nu, refs, scene, kind, solids, warm) in enumerate(LEVEL_CASES):
        probe = R.run(R.RefCase(dim=dim, grid_size=gs, non_uniform=nu, refine=refs))
        shape = (probe["Ny"], probe["Nx"], probe["Nz"])
        rng = np.random.default_rng(4000 + q)
        t = np.zeros(shape, np.uint8)
        ring = ring_mask(shape, dim)
        t[ring] = INFLOW if kind == "inflow" else SOLID
        if kind == "outlet":
            t[:, -1][ring[:, -1]] = EMPTY
        if solids:
            inner = ~ring
            t[inner & (rng.random(shape) < 0.12)] = SOLID
        inner_t = t[1:-1, 1:-1, 1:-1] if dim == 3 else t[1:-1, 1:-1, :]
        n = int((inner_t == FLUID).sum())
        x = rng.standard_normal(n) + 2.5
        p_prev = rng.standard_normal(t.size) + 1.0 if warm == "random" else np.zeros(t.size)
        out = R.run(R.RefCase(dim=dim, grid_size=gs, non_uniform=nu, refine=refs, types=t, compat=True,
                              extras={"scene": scene, "t": 0.3, "x": x, "p_prev": p_prev}), threads=1)
        wz = out["wz"] if dim == 3 else np.array([out["dh"]])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            dim=dim, non_uniform=int(out["non_uniform"]), dh=out["dh"], wx=out["wx"], wy=out["wy"], wz=wz, types=t,
            singular=int(out["singular"]), order=out["order"], x=x, t=0.3, scene=scene,
            inflow_level=out["inflow_level"], cx=out["cx"], cy=out["cy"], cz=out["cz"], x_levelled=out["x_levelled"],
            p_prev=p_prev, warm_built=int(out["warm_built"]), warm_guess=out["warm_guess"])
        print(f"{name}: n={out['n']} singular={out['singular']} levels={int(np.isfinite(out['inflow_level']).sum())} "
              f"warm={out['warm_built']}")


if __name__ == "__main__":
    if "--level" in sys.argv:
        main_level()
        raise SystemExit(0)
    if "--floor" in sys.argv:
        main_floor()
        raise SystemExit(0)
    if "--fields" not in sys.argv:
        main()
    main_fields()
    main_floor()
    main_level()
