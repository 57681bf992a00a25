"""Write golden trajectory fixtures by running the reference step on the CPU.

Only the composed multi-step trajectory of the reference's ``one_step``
(``step_time.py:1574-1637``: radial loop, CFL, IMEX weights, fused lm loop,
insulating B update) cannot be pinned by closed forms or library routines,
so this script imports the reference package from ``/root/reference/src``
(pure Python, FP64, ``MAGIC_DEVICE=cpu``) in a child process per
configuration and stores small ``.npz`` vectors here.  It is run by hand in
the development container; tests only read the ``.npz`` files.

Cases (each: ``initialize_fields`` → optional seeded noise on s →
``setup_initial_state`` → ``initialize_dt`` → ``one_step`` × n):

* ``bench16``: dynamo benchmark l_max=16, n_r_max=33, no noise, 20 steps;
  energies every step, fields at steps 1 and 20, explicit terms after the
  first radial loop.
* ``c1``: l_max=32, n_r_max=33, E=1e-3, Ra=1e5, Pr=1, Pm=5, dt=1e-4,
  noise 1e-5 (seed 20261202), 100 steps; energies every step, fields at
  step 100.

Usage: ``python tests/golden/make_golden.py``
"""

from __future__ import annotations

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"

CASES = {
    "bench16": dict(cfg=dict(l_max=16, n_r_max=33), dt=1e-4, steps=20, noise=None,
                    field_steps=[1, 20], expl_step1=True),
    "c1": dict(cfg=dict(l_max=32, n_r_max=33, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0,
                        radratio=0.35, dtmax=1e-4), dt=1e-4, steps=100,
               noise=dict(seed=20261202, amp=1e-5), field_steps=[100], expl_step1=False),
}

CHILD = r"""
import json, os, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]
sys.path.insert(0, sys.argv[3]); sys.path.insert(0, sys.argv[4])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields, dt_fields, step_time, output
from magic_torch.init_fields import initialize_fields
initialize_fields()
if case["noise"]:
    from oracle.params import Params
    from oracle import horizontal, radial
    from oracle.dynamo import seeded_noise
    p = Params(case["cfg"])
    nz = seeded_noise(p, horizontal.lm_maps(p), radial.radial_functions(p),
                      case["noise"]["seed"], case["noise"]["amp"])
    fields.s_LMloc += torch.from_numpy(nz)
step_time.setup_initial_state()
step_time.initialize_dt(case["dt"])
names = ["s", "z", "w", "p", "b", "aj"]
res = {}
en = []
dts = []
for n in range(1, case["steps"] + 1):
    dt = step_time.one_step(n, case["dt"])
    dts.append(dt)
    ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
    em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
    en.append([ek.e_p, ek.e_t, em.e_p, em.e_t])
    if n in case["field_steps"]:
        for k in names:
            res[f"{k}_step{n}"] = getattr(fields, k + "_LMloc").numpy().copy()
    if n == 1 and case["expl_step1"]:
        for k, ta in (("s", dt_fields.dsdt), ("z", dt_fields.dzdt), ("w", dt_fields.dwdt),
                      ("p", dt_fields.dpdt), ("b", dt_fields.dbdt), ("j", dt_fields.djdt)):
            # after the lm loop the step-1 explicit term sits in slot 1
            res[f"expl_{k}_step1"] = ta.expl[:, :, 1].numpy().copy()
res["energies"] = np.array(en)
res["dt"] = np.array(dts)
np.savez(out, **res)
"""


def main():
    env = dict(os.environ, MAGIC_DEVICE="cpu", PYTHONDONTWRITEBYTECODE="1",
               TORCH_COMPILE_DISABLE="1", TORCHDYNAMO_DISABLE="1")
    for name, case in CASES.items():
        out = os.path.join(HERE, f"traj_{name}.npz")
        subprocess.run([sys.executable, "-c", CHILD, json.dumps(case), out, REF_SRC, REPO],
                       check=True, env=env)
        meta = dict(case, source="rvilim/torch-magic src/magic_torch step_time.one_step (FP64 CPU)")
        with open(os.path.join(HERE, f"traj_{name}.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print("wrote", out)


if __name__ == "__main__":
    main()
