"""Writes the golden vectors under tests/golden/ from the CPU oracle.

The reference ships no golden vectors for this path and cannot be built or run here, so these
fixtures are the oracle's outputs on small seeded inputs, committed as a regression lock: the
CPU tests check that the oracle still reproduces them, the GPU tests that libcn_hip.so does.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "cressim-neo_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cn_oracle  # noqa: E402
import cnhip  # noqa: E402
from cnhip import scenes  # noqa: E402


def params(**kw):
    return cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64, **kw)


def broadphase():
    """Keys, cells, 1-bit sorted pairs and the cell-range table of 600 points, 50 on knife edges."""
    rng = np.random.default_rng(11)
    pts = np.zeros((600, 4), np.float32)
    pts[:, :3] = rng.uniform(-1.0, 1.0, (600, 3))
    pts[:50, :3] = np.round(pts[:50, :3] / 0.1) * np.float32(0.1)
    keys, cells = cn_oracle.broadphase_entries(pts, 0.1)
    sk, si = cn_oracle.radix_sort_1bit(keys, np.arange(600, dtype=np.uint32))
    cap = cn_oracle.next_pow2(2 * 600)
    tbl = cn_oracle.cell_ranges(sk, cap)
    np.savez_compressed(os.path.join(HERE, "broadphase_600.npz"), pos=pts, cell=np.float32(0.1), keys=keys,
                        cells=cells, sorted_keys=sk, sorted_idx=si, table=tbl)


def dam_break_steps():
    """A 10^3 jittered dam-break block (config [0] geometry, r = 0.03): iteration-0 neighbor lists of
    the first substep, then the state after each of 3 steps of 4 substeps x 4 iterations."""
    sc = scenes.dam_break_64k(side=10)
    p = params()
    osc = cn_oracle.OracleScene(sc)
    nb = cn_oracle.substep_neighbors(osc, p, float(np.max(sc.fluid_mats["smoothing_radius"])))
    out = dict(pos0=sc.pos.copy(), vel0=sc.vel.copy(), counts=nb["counts"], lists=nb["lists"][:, :48])
    for s in range(3):
        cn_oracle.step(osc, p)
        out[f"pos{s + 1}"] = osc.pos.copy()
        out[f"vel{s + 1}"] = osc.vel.copy()
    np.savez_compressed(os.path.join(HERE, "dam_break_1000.npz"), **out)


def soft_cloth_steps():
    """A 5^3 tet block and a 12 x 12 cloth over a 6^3 fluid block: state after 2 steps."""
    sc = scenes.compose([scenes.soft_block_part(5, 0.05), scenes.cloth_part(12, 0.03),
                         scenes.fluid_part(6, 0.05, 0.03, origin=(1.0, 0.05, 1.0), seed=3, jitter=0.005)],
                        cnhip.fluid_material(0.03), scenes.container((0, 0, 0), (4, 2, 2)), "golden_soft")
    p = params()
    osc = cn_oracle.OracleScene(sc)
    out = dict(pos0=sc.pos.copy(), vel0=sc.vel.copy())
    for s in range(2):
        cn_oracle.step(osc, p)
        out[f"pos{s + 1}"] = osc.pos.copy()
        out[f"vel{s + 1}"] = osc.vel.copy()
    np.savez_compressed(os.path.join(HERE, "soft_cloth_fluid.npz"), **out)


if __name__ == "__main__":
    broadphase()
    dam_break_steps()
    soft_cloth_steps()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
