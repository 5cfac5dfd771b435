"""Write the golden fixtures of this directory from EMopt's own CW-FDTD kernel.

The reference kernel is src/fdtd.cpp compiled from its own sources by
``make -C oracle ref`` into oracle/_ref/libfdtd_ref.so; it is driven through
the reference's FDTD_* C ABI by the oracle driver's single-rank restatement of
emopt/fdtd.py (__init__, sources, the GhostComm face copy, __solve).

Fixtures (inputs are regenerated from the recorded seeds/parameters by
tests/cases.py; only outputs are stored):

* steps.npz   time-domain Ex..Hz (interior of the padded arrays) after given
              step counts, random materials / complex J+M blocks, every bc of
              {000, 0H0, E00, HE0}, real and complex eps, CPML widths 0/3/4,
              more source blocks than the GPU kernel carries inline
* solve_strip_24x20x16.npz, solve_strip_40x32x24.npz  full converged solves
              of the strip waveguide: n0, Tn, the convergence history and the
              six recovered complex components
* kats.npz    scalar known answers of calc_phase_3T / calc_amplitude_3T and
              src_func_t

Run:  make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import driver as D  # noqa: E402
from tests.cases import random_case, run_oracle, strip_case  # noqa: E402
from tests.test_oracle import run_small_steps  # noqa: E402

STEP_CASES = []
for bc in ("000", "0H0", "E00", "HE0"):
    for ceps in (0, 1):
        STEP_CASES.append(dict(Nx=12, Ny=10, Nz=8, seed=10 + len(STEP_CASES), ceps=ceps, nsrc=2,
                               w=(3, 3, 4, 3, 3, 4), bc=bc, steps=(1, 2, 10, 50)))
STEP_CASES.append(dict(Nx=24, Ny=20, Nz=16, seed=40, ceps=0, nsrc=1, w=(4, 3, 3, 4, 4, 3),
                       bc="000", steps=(1, 2, 10, 50)))
STEP_CASES.append(dict(Nx=12, Ny=10, Nz=8, seed=41, ceps=0, nsrc=6, w=(0, 0, 0, 0, 0, 0),
                       bc="0H0", steps=(20,)))
for q, bc in enumerate(("0H0", "E00", "HE0")):
    STEP_CASES.append(dict(Nx=24, Ny=20, Nz=16, seed=42 + q, ceps=q % 2, nsrc=2,
                           w=(4, 3, 3, 4, 4, 3), bc=bc, steps=(1, 2, 10, 50)))


def main():
    if not os.path.exists(D.REF_LIB):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref = D.load(D.REF_LIB)
    out = {}
    for q, c in enumerate(STEP_CASES):
        case = random_case(c["Nx"], c["Ny"], c["Nz"], w=c["w"], seed=c["seed"],
                           ceps=bool(c["ceps"]), bc=c["bc"], nsrc=c["nsrc"])
        o = run_small_steps(ref, case, 0)
        fields, done = [], 0
        for n_target in c["steps"]:
            while done < n_target:
                o._step(done)
                done += 1
            fields.append(np.stack([o.padded(k)[1:-1, 1:-1, 1:-1].copy() for k in range(6)]))
        key = "case%02d" % q
        out[key + "_meta"] = np.array([c["Nx"], c["Ny"], c["Nz"], c["seed"], c["ceps"], c["nsrc"],
                                       *c["w"], *[ord(ch) for ch in c["bc"]]], dtype=np.int64)
        out[key + "_steps"] = np.array(c["steps"], dtype=np.int64)
        out[key + "_fields"] = np.stack(fields)
    np.savez_compressed(os.path.join(HERE, "steps.npz"), **out)

    for dims, w in (((24, 20, 16), 4), ((40, 32, 24), 5)):
        r = run_oracle(strip_case(*dims, w=w), lib="ref")
        np.savez_compressed(os.path.join(HERE, "solve_strip_%dx%dx%d.npz" % dims), n0=r["n0"],
                            Tn=r["Tn"], history=r["history"], fields=np.stack(r["fields"]))

    rng = np.random.default_rng(99)
    ph_in, ph_out, am_in, am_out = [], [], [], []
    for _ in range(64):
        t = np.sort(rng.uniform(0, 200, 3))
        f = rng.standard_normal(3) * rng.choice([1e-6, 1.0, 1e3])
        if rng.random() < 0.1:
            f[:] = f[0]
        p = ref.FDTD_calc_phase_3T(*t, *f)
        a = ref.FDTD_calc_amplitude_3T(*t, *f, p)
        ph_in.append([*t, *f]); ph_out.append(p)
        am_in.append([*t, *f, p]); am_out.append(a)
    h = ref.FDTD_new()
    T, m = 31.4159, 1e-4
    ref.FDTD_set_source_properties(h, T, m)
    src_in, src_out = [], []
    for t in np.linspace(0, 2 * T, 41):
        ph = rng.uniform(-np.pi, np.pi)
        src_in.append([t, ph]); src_out.append(ref.FDTD_src_func_t(h, 0, t, ph))
    np.savez_compressed(os.path.join(HERE, "kats.npz"), phase3_in=np.array(ph_in),
                        phase3_out=np.array(ph_out), amp3_in=np.array(am_in),
                        amp3_out=np.array(am_out), src_T=T, src_min=m,
                        src_in=np.array(src_in), src_out=np.array(src_out))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
