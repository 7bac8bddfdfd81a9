"""Writes the golden fixtures of this directory (run from the repository root:
python tests/golden/make_golden.py).

* onsager_table.json - -beta f of the 2D Ising model from Onsager's closed
  form, evaluated with mpmath at 30 digits on a beta grid (independent of the
  scipy quadrature in oracle/onsager.py).
* reference_values.json - the numbers the reference itself publishes for this
  path (README.md and test/*.jl), with their file:line.
* oracle_norms.json - per-step norm vectors of the FP64 oracle on fixed
  models and configurations, so the GPU parity tests compare against stored
  vectors and the CPU suite notices any drift of the oracle.
* oracle_extensions.json - the same for the schemes and observables beyond
  TRG / BTRG / HOTRG: ATRG and HOTRG_3D norm vectors, ImpurityTRG (pivoted
  SVD gauge) and ImpurityHOTRG rows, and cft_data / ground_state_degeneracy /
  gu_wen_ratio / central_charge after fixed runs
  (python tests/golden/make_golden.py extensions writes only this file).
* oracle_configs.json - norm vectors at the benchmark configurations' sizes
  (BTRG chi=32 x 30 at three betas of the seeded beta scan; TRG and HOTRG
  on the Trivial Ising tensor at beta_c, chi 32 and 48, 30 steps; BTRG Z2
  Ising and Z3 Potts at chi 64, 30 steps; the first 16 runs of the seeded
  512-run scan with chi in {16, 24, 32, 48}), each with its truncation log:
  per factorization the matrix size, the largest value, the smallest kept and
  the largest dropped value and the 2-norm of the dropped ones
  (python tests/golden/make_golden.py configs [name ...] writes only this
  file; names select cases and merge into the existing file), and the
  oracle's own roundoff sensitivity per norm ("sens": the same run with a
  seeded 1e-15 perturbation of every factorized matrix;
  python tests/golden/make_golden.py configs-sens [name ...]).
"""
from __future__ import annotations

import json
import os
import sys

import mpmath

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import models as om        # noqa: E402
from oracle import schemes as osch     # noqa: E402

mpmath.mp.dps = 30


def onsager_mp(beta):
    b = mpmath.mpf(beta)
    k = 1 / mpmath.sinh(2 * b) ** 2
    c2 = mpmath.cosh(2 * b) ** 2
    f = lambda t: mpmath.log(c2 + mpmath.sqrt(1 + k * k - 2 * k * mpmath.cos(2 * t)) / k)
    val = mpmath.quad(f, [0, mpmath.pi / 4, mpmath.pi / 2, 3 * mpmath.pi / 4, mpmath.pi])
    return -(mpmath.log(2) / 2 + val / (2 * mpmath.pi)) / b


# (name, scheme, model, chi, maxiter)
NORM_CASES = [
    ("btrg_z2_ising_betac_chi16_25", "BTRG", ("ising_z2", None), 16, 25),
    ("trg_z2_ising_betac_chi16_12", "TRG", ("ising_z2", None), 16, 12),
    ("hotrg_z2_ising_betac_chi8_10", "HOTRG", ("ising_z2", None), 8, 10),
    ("btrg_trivial_ising_b0.40_chi24_10", "BTRG", ("ising_trivial", 0.40), 24, 10),
    ("btrg_z3_potts_betac_chi12_10", "BTRG", ("potts_z3", None), 12, 10),
    ("trg_z6_clock_b1.1_chi12_8", "TRG", ("clock_z6", 1.1), 12, 8),
]


def model(spec):
    name, beta = spec
    if name == "ising_z2":
        return om.classical_ising_z2(om.ising_betac_mp if beta is None else beta)
    if name == "ising_trivial":
        return om.classical_ising_trivial(beta)
    if name == "potts_z3":
        return om.classical_potts_zn(3, om.potts_betac(3) if beta is None else beta)
    if name == "clock_z6":
        return om.classical_clock_zn(6, beta)
    raise ValueError(name)


def main():
    grid = [round(0.30 + 0.02 * i, 2) for i in range(16)] + [float(om.ising_betac)]
    table = {"note": "-beta f from Onsager's closed form, mpmath 30 digits",
             "beta": grid, "f": [float(onsager_mp(b)) for b in grid]}
    with open(os.path.join(HERE, "onsager_table.json"), "w") as f:
        json.dump(table, f, indent=1)

    ref = {
        "readme_btrg_z2_chi16_maxiter25_f": {
            "value": "-2.1096504926141826902647832", "source": "README.md:87"},
        "readme_btrg_rel_err_vs_onsager": {"value": 3.1e-7, "source": "README.md:91-94"},
        "f_onsager_betac": {"value": "-2.1096511446082074596677792835110847808254932754354",
                            "source": "src/models/ising.jl:2"},
        "test_gates_vs_onsager": {
            "TRG_chi24_maxiter25": {"rtol": 2e-6, "source": "test/schemes.jl:20-23"},
            "BTRG_chi24_maxiter25": {"rtol": 6e-8, "source": "test/schemes.jl:59-62"},
            "HOTRG_chi16_maxiter25": {"rtol": 6e-7, "scalefactor": 4,
                                      "source": "test/schemes.jl:96-99"}},
    }
    with open(os.path.join(HERE, "reference_values.json"), "w") as f:
        json.dump(ref, f, indent=1)

    out = {}
    for name, kind, spec, chi, n in NORM_CASES:
        sc = {"TRG": osch.TRG, "BTRG": osch.BTRG, "HOTRG": osch.HOTRG}[kind](model(spec))
        d = osch.run(sc, chi, n)
        out[name] = {"scheme": kind, "model": spec[0], "beta": spec[1], "chi": chi,
                     "maxiter": n, "norms": [float(x) for x in d]}
        print(name, len(d), flush=True)
    with open(os.path.join(HERE, "oracle_norms.json"), "w") as f:
        json.dump(out, f, indent=1)


def extensions():
    from oracle import observables as ob
    out = {}
    sc = osch.ATRG(om.classical_ising_z2(om.ising_betac_mp))
    out["atrg_z2_ising_betac_chi12_8"] = {"chi": 12, "maxiter": 8,
                                          "norms": [float(x) for x in osch.run(sc, 12, 8)]}
    sc = osch.HOTRG_3D(om.classical_ising_3d_z2(om.ising_betac_3d))
    out["hotrg3d_z2_ising3d_betac_chi6_3"] = {"chi": 6, "maxiter": 3,
                                              "norms": [float(x) for x in osch.run(sc, 6, 3)]}
    T, Ti = om.classical_ising_trivial(0.44, 0.1), om.classical_ising_impurity(0.44, 0.1)
    d = osch.run(osch.ImpurityTRG(T, Ti, T, T, T, gauge="pivoted"), 8, 6)
    out["impurity_trg_b0.44_h0.1_chi8_6"] = {"beta": 0.44, "h": 0.1, "chi": 8, "maxiter": 6,
                                             "rows": [[float(v) for v in r] for r in d]}
    T, Ti = om.classical_ising_trivial(0.5, 0.1), om.classical_ising_impurity(0.5, 0.1)
    d = osch.run(osch.ImpurityHOTRG(T, Ti, Ti, T), 8, 4)
    out["impurity_hotrg_b0.5_h0.1_chi8_4"] = {"beta": 0.5, "h": 0.1, "chi": 8, "maxiter": 4,
                                              "rows": [[float(v) for v in r] for r in d]}
    for kind in ("TRG", "BTRG"):
        sc = {"TRG": osch.TRG, "BTRG": osch.BTRG}[kind](om.classical_ising_z2(om.ising_betac_mp))
        d = osch.run(sc, 16, 6)
        x1, x2 = ob.gu_wen_ratio(sc)
        out[f"observables_{kind.lower()}_z2_ising_betac_chi16_6"] = {
            "scheme": kind, "chi": 16, "maxiter": 6,
            "cft_data_first6": [float(v) for v in ob.cft_data(sc)[:6]],
            "ground_state_degeneracy": ob.ground_state_degeneracy(sc),
            "gu_wen_ratio": [x1, x2],
            "central_charge_n_last": ob.central_charge(sc, d[-1]),
            "n_last": float(d[-1])}
    with open(os.path.join(HERE, "oracle_extensions.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(out), "cases")


# seeds of the benchmark workloads: the headline beta scan and the 512-run
# mixed-chi scan draw from numpy default_rng(SCAN_SEED) and (SCAN_SEED + 4)
SCAN_SEED = 20260512


def config_cases():
    """(name, scheme, model spec, chi, maxiter) at the benchmark sizes."""
    import numpy as np
    cases = []
    betas = np.random.default_rng(SCAN_SEED).uniform(0.35, 0.55, 128)
    for i in range(3):
        cases.append((f"scan_btrg_trivial_ising_run{i}_chi32_30", "BTRG",
                      ("ising_trivial", float(betas[i])), 32, 30))
    for kind in ("TRG", "HOTRG"):
        for chi in (32, 48):
            cases.append((f"c2_{kind.lower()}_trivial_ising_betac_chi{chi}_30", kind,
                          ("ising_trivial_betac", None), chi, 30))
    cases.append(("c3_btrg_z2_ising_betac_chi64_30", "BTRG", ("ising_z2", None), 64, 30))
    cases.append(("c3_btrg_z3_potts_betac_chi64_30", "BTRG", ("potts_z3", None), 64, 30))
    rng = np.random.default_rng(SCAN_SEED + 4)
    b4 = rng.uniform(0.35, 0.55, 512)
    c4 = rng.choice([16, 24, 32, 48], 512)
    for i in range(16):
        cases.append((f"c4_btrg_trivial_ising_run{i:02d}_chi{int(c4[i])}_30", "BTRG",
                      ("ising_trivial", float(b4[i])), int(c4[i]), 30))
    return cases


def config_model(spec):
    name, beta = spec
    if name == "ising_trivial_betac":
        return om.classical_ising_trivial(om.ising_betac_mp)
    return model(spec)


def _merge_into(path, name, entry):
    """Read-merge-write of one case under an exclusive lock (several
    generator processes may run side by side on disjoint cases)."""
    import fcntl
    with open(path + ".lock", "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        out = {}
        if os.path.exists(path):
            with open(path) as f:
                out = json.load(f)
        out[name] = entry
        with open(path + ".tmp", "w") as f:
            json.dump(dict(sorted(out.items())), f, indent=1)
        os.replace(path + ".tmp", path)


def configs(names=()):
    import time
    path = os.path.join(HERE, "oracle_configs.json")
    for name, kind, spec, chi, n in config_cases():
        if names and name not in names:
            continue
        t0 = time.time()
        sc = {"TRG": osch.TRG, "BTRG": osch.BTRG, "HOTRG": osch.HOTRG}[kind](config_model(spec))
        d = osch.run(sc, chi, n)
        beta = float(om.ising_betac if spec[1] is None and spec[0] != "potts_z3"
                     else om.potts_betac(3) if spec[1] is None else spec[1])
        _merge_into(path, name, {
            "scheme": kind, "model": spec[0], "beta": spec[1], "chi": chi, "maxiter": n,
            "norms": [float(x) for x in d],
            "f": float(osch.free_energy(d, beta, sc.scalefactor)),
            "cuts": [[int(c.n), c.top, c.kept, c.first_dropped, c.err] for c in sc.log]})
        print(name, f"{time.time() - t0:.1f} s", flush=True)


def _perturbed_run(kind, spec, chi, n, seed=7, rel=1e-15):
    """The same oracle run with every matrix the truncated factorizations
    see perturbed by seeded noise of `rel` times its largest entry (the size
    of the roundoff two correct implementations differ by)."""
    import numpy as np
    from oracle import ztensor as zt
    rng = np.random.default_rng(seed)
    svd0, eigh0 = osch.svd_trunc, osch.eigh_trunc

    def noisy(t):
        d = t.data + rng.standard_normal(t.data.shape) * rel * float(np.abs(t.data).max())
        # keep the symmetry: entries that violate charge conservation stay 0
        tot = np.zeros(t.data.shape, dtype=np.int64)
        for leg, x in enumerate(t.q):
            shp = [1] * t.data.ndim
            shp[leg] = len(x)
            tot = tot + x.reshape(shp)
        d[(tot % t.N) != 0] = 0.0
        return zt.ZT(d, t.q, t.N)

    def svd_p(t, ncod, c, gauge="lapack"):
        return svd0(noisy(t), ncod, c, gauge)

    def eigh_p(t, ncod, c):
        return eigh0(zt.project_hermitian(noisy(t), ncod), ncod, c)

    osch.svd_trunc, osch.eigh_trunc = svd_p, eigh_p
    try:
        sc = {"TRG": osch.TRG, "BTRG": osch.BTRG, "HOTRG": osch.HOTRG}[kind](config_model(spec))
        return osch.run(sc, chi, n)
    finally:
        osch.svd_trunc, osch.eigh_trunc = svd0, eigh0


def configs_sensitivity(names=()):
    """Adds to each case of oracle_configs.json "sens": per norm, the
    relative change under a 1e-15 perturbation of every factorized matrix
    (_perturbed_run).  Steps where it exceeds ~1e-12 are not determined to
    1e-10 by the tensor: the critical RG flow amplifies roundoff there, or
    a truncation splits a near-degenerate multiplet."""
    import time
    path = os.path.join(HERE, "oracle_configs.json")
    for name, kind, spec, chi, n in config_cases():
        if names and name not in names:
            continue
        with open(path) as f:
            cur = json.load(f).get(name)
        if cur is None:
            continue
        t0 = time.time()
        d = _perturbed_run(kind, spec, chi, n)
        cur["sens"] = [abs(x / y - 1) for x, y in zip(d, cur["norms"])]
        _merge_into(path, name, cur)
        print(name, "sens", f"{max(cur['sens']):.2e}", f"{time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["extensions"]:
        extensions()
    elif sys.argv[1:2] == ["configs"]:
        configs(tuple(sys.argv[2:]))
    elif sys.argv[1:2] == ["configs-sens"]:
        configs_sensitivity(tuple(sys.argv[2:]))
    else:
        main()
