"""Config 3, the coupled heterogeneous system, assembled from the device
producers and solved through the C ABI: affine bodies (12-DoF blocks,
ortho potential, 3x3 blocks in a 4x4 grid) + a Baraff-Witkin cloth with
dihedral bending + vertex-triangle contacts on the cloth, one global system
with two local preconditioners (ABDDiagPreconditioner on the bodies' range,
MAS on the cloth's range, its partition from mi_mesh_partition and its
hierarchy built on the device).  At a reduced size the GPU solve is checked
against the oracle's PCG on the same triplets; at config 3's size (10,000
bodies, 256 x 256 cloth, 500k contacts) by its true residual."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.abi import check
from libuipc_amd.device import LinearSystem, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def coupled(ctx, n_bodies, grid, n_contacts, seed):
    a = synth.abd_inputs(n_bodies, seed=seed)
    c = synth.cloth_inputs(grid, seed=seed)
    hg = synth.hinges_of(c.tris)
    st, g, kap = synth.cloth_contacts(grid, n_contacts, seed=seed)
    Bb, V = 4 * n_bodies, grid * grid
    B = Bb + V
    nt, nh, nc = len(c.tris), len(hg), len(st)
    T = 16 * n_bodies + V + 6 * nt + 10 * nh + 10 * nc
    rows, cols, vals = ctx.alloc(T, np.int32), ctx.alloc(T, np.int32), ctx.alloc((T, 9), np.float64)
    r0, c0, v0 = rows.ptr.value, cols.ptr.value, vals.ptr.value
    dev = {k: ctx.to_device(v) for k, v in dict(q=a.q, dy=a.dyadic, vol=a.volume, ka=a.kappa, tris=c.tris,
                                                 X=c.rest_positions, x=c.positions, hg=hg, st=st, g=g, kap=kap,
                                                 m=np.full(V, 200.0 * 2e-3 / V)).items()}
    lib, off = ctx.lib, 0
    check(lib.mi_produce_abd(ctx.h, n_bodies, dev["q"].ptr, dev["dy"].ptr, dev["vol"].ptr, dev["ka"].ptr, None, a.dt,
                             0, r0, c0, v0, None))
    off += 16 * n_bodies
    check(lib.mi_produce_kinetic(ctx.h, V, dev["m"].ptr, Bb, r0 + 4 * off, c0 + 4 * off, v0 + 72 * off, None))
    off += V
    check(lib.mi_produce_bw_shell(ctx.h, nt, dev["tris"].ptr, dev["X"].ptr, dev["x"].ptr, c.stretch, c.shear,
                                  c.strain_rate, c.dt, Bb, r0 + 4 * off, c0 + 4 * off, v0 + 72 * off, None))
    off += 6 * nt
    check(lib.mi_produce_bending(ctx.h, nh, dev["hg"].ptr, dev["X"].ptr, dev["x"].ptr, c.bending, c.dt, Bb,
                                 r0 + 4 * off, c0 + 4 * off, v0 + 72 * off, None))
    off += 10 * nh
    check(lib.mi_produce_contact(ctx.h, nc, dev["st"].ptr, dev["g"].ptr, dev["kap"].ptr, c.dt, Bb, r0 + 4 * off,
                                 c0 + 4 * off, v0 + 72 * off, None))
    ctx.synchronize()
    parts, _ = ctx.mesh_partition(V, 16, tris=c.tris)
    return dict(B=B, Bb=Bb, V=V, rows=rows, cols=cols, vals=vals, parts=parts, tris=c.tris, keep=dev)


def solve(ctx, s, rhs, tol):
    ls = LinearSystem(ctx, s["B"])
    ls.assemble(s["rows"], s["cols"], s["vals"])
    ls.add_preconditioner(abi.MI_PRECOND_DIAG12, 0, 3 * s["Bb"])
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 3 * s["Bb"], 3 * s["V"], part_ids=s["parts"], tris=s["tris"],
                          build=abi.MI_MAS_BUILD_DEVICE)
    ls.assemble_preconditioners()
    x = ctx.alloc(3 * s["B"], np.float64)
    st = ls.solve(ctx.to_device(rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=tol, true_residual=1)
    ls.close()
    return x.numpy(), st


def test_coupled_small_against_oracle(ctx):
    s = coupled(ctx, 50, 24, 2000, seed=3)
    rhs = np.random.default_rng(1).uniform(-1, 1, 3 * s["B"])
    x, st = solve(ctx, s, rhs, 1e-6)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-6
    r, c, v = s["rows"].numpy(), s["cols"].numpy(), s["vals"].numpy()
    orr, oc, ov, _ = O.convert_sym(s["B"], r, c, v)
    m = O.Mas(s["V"], s["parts"], tris=s["tris"])
    cloth = slice(3 * s["Bb"], 3 * s["B"])
    # the cloth's own blocks for the MAS (the subsystem range of the global matrix)
    keep = (orr >= s["Bb"]) & (oc >= s["Bb"])
    m.assemble(orr[keep] - s["Bb"], oc[keep] - s["Bb"], ov[keep])
    pc = [dict(kind=2, blk_off=0, blk_cnt=s["Bb"], data=O.diag12(orr, oc, ov, 0, s["Bb"] // 4)),
          dict(kind=3, blk_off=s["Bb"], blk_cnt=s["V"], mas=m)]
    xo, ro = O.pcg(s["B"], orr, oc, ov, rhs, pc, criterion=1, tol=1e-6)
    assert ro["converged"] == 1
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(2, ro["iters_exact"] // 10)
    assert rel_err(x, xo) < 1e-3
    assert rel_err(x[cloth], xo[cloth]) < 1e-3


def test_coupled_config3_size(ctx):
    s = coupled(ctx, 10_000, 256, 500_000, seed=4)
    rhs = np.random.default_rng(2).uniform(-1, 1, 3 * s["B"])
    x, st = solve(ctx, s, rhs, 1e-4)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-4
    print(f"\n[coupled] 10,000 bodies + 256^2 cloth + {len(s['keep']['st'].numpy())} contacts: B={s['B']}, "
          f"{st['iters_exact']} PCG iterations to 1e-4", flush=True)
