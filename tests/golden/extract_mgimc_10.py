"""Writes tests/golden/mgimc_10_powerlaw.json from the reference's own log of
its 10-group power-law Marshak-wave run
(results/marshak_wave_powerlaw/imc_10g_publication/mgimc_10.out, written by
MG_IMC/problems/test_marshak_wave_multigroup_powerlaw.py --mode publication).

Kept as data: the Planck emission fractions chi_g and the geometric-mean
group opacities printed at T = 0.05 keV (lines 15-25; computed by the
reference with its planck_integrals module), the per-step table (step, t,
N, E_mat, E_rad) and the summary rows (t, N, total, internal and radiation
energy, boundary emission and outgoing energy of the step, source emission,
residual) of the first steps.  Runs only where the reference is present."""
import json
import os
import re
import sys

SRC = "/root/reference/results/marshak_wave_powerlaw/imc_10g_publication/mgimc_10.out"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mgimc_10_powerlaw.json")
MAX_STEP = 100


def main():
    lines = open(SRC).read().splitlines()
    chi, sig, lo, hi = [], [], [], []
    pat = re.compile(r"Group\s+(\d+) \[\s*([\d.]+),\s*([\d.]+)\] keV: chi = ([\d.]+), sigma_a = ([\d.e+-]+) cm")
    for ln in lines:
        m = pat.search(ln)
        if m:
            lo.append(float(m.group(2))); hi.append(float(m.group(3)))
            chi.append(float(m.group(4))); sig.append(float(m.group(5)))
    steps, summary = [], []
    for ln in lines:
        f = ln.split()
        if len(f) == 8 and f[0].isdigit() and int(f[0]) <= MAX_STEP:
            steps.append([int(f[0]), float(f[1]), int(f[3]), float(f[4]), float(f[5])])
        elif len(f) == 9 and "\t" in ln and re.match(r"^\d\.\d+e[-+]\d+$", f[0]):
            row = [float(f[0]), int(f[1])] + [float(x) for x in f[2:]]
            if row[0] <= MAX_STEP * 0.01 + 1e-9:
                summary.append(row)
    data = {
        "source": "results/marshak_wave_powerlaw/imc_10g_publication/mgimc_10.out",
        "T_chi": 0.05,
        "group_lo_printed": lo, "group_hi_printed": hi,
        "chi": chi, "sigma_a": sig,
        "steps_columns": ["step", "t_ns", "N_particles", "E_mat", "E_rad"],
        "steps": steps,
        "summary_columns": ["t_ns", "N_particles", "total_energy", "internal_energy", "radiation_energy",
                            "boundary_emission", "boundary_outgoing", "source_emission", "energy_residual"],
        "summary": summary,
    }
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=1)
    print(f"wrote {OUT}: {len(chi)} groups, {len(steps)} steps, {len(summary)} summary rows")


if __name__ == "__main__":
    sys.exit(main())
