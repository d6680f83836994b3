"""Writes the golden fixtures under tests/golden/ from data files that the
reference repository holds (no reference code is imported or run).

Sources (reference repository, read as data):
  notebooks/quick_single_run/data/{run.ini,events.log,thermodynamics.csv,observations.csv}
      produced by the reference's own binary: N=500, rho=0.8, T=0.5, rc=2.5,
      dt=0.005, seed 117208264.
  notebooks/quick_sweep_run/sweep.ini and data/T_*/d_*/{run.ini,events.log}
      the 540-run demo sweep.

Outputs (small; the reference is GPL-3.0, the data are attributed to it):
  quick_single_run/run.ini, events.log, observations.csv  (verbatim)
  quick_single_run/thermodynamics_head.csv   header + rows 0..199 (verbatim lines)
  quick_single_run/thermodynamics_every50.csv header + every 50th row to 1599
  quick_sweep_run/sweep.ini                   (verbatim)
  quick_sweep_run/runs.csv  one line per run directory, in T-major order:
      directory, temperature, density, random_seed, equilibration name,
      observation name, total time steps (last events.log step), status

Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""

import configparser
import pathlib
import shutil
import sys

HERE = pathlib.Path(__file__).resolve().parent


def main(ref="/root/reference"):
    ref = pathlib.Path(ref)
    src = ref / "notebooks/quick_single_run/data"
    dst = HERE / "quick_single_run"
    dst.mkdir(exist_ok=True)
    for name in ["run.ini", "events.log", "observations.csv"]:
        shutil.copyfile(src / name, dst / name)
    lines = (src / "thermodynamics.csv").read_text().splitlines(keepends=True)
    (dst / "thermodynamics_head.csv").write_text("".join(lines[:201]))
    every = [lines[0]] + [lines[1 + k] for k in range(0, 1600, 50)]
    (dst / "thermodynamics_every50.csv").write_text("".join(every))

    sw = ref / "notebooks/quick_sweep_run"
    sdst = HERE / "quick_sweep_run"
    sdst.mkdir(exist_ok=True)
    shutil.copyfile(sw / "sweep.ini", sdst / "sweep.ini")
    out = ["directory,temperature,density,random_seed,equilibration_name,observation_name,"
           "total_time_steps,status\n"]
    tdirs = sorted((sw / "data").glob("T_*"), key=lambda p: float(p.name[2:]))
    for tdir in tdirs:
        for ddir in sorted(tdir.glob("d_*"), key=lambda p: float(p.name[2:])):
            cp = configparser.ConfigParser()
            cp.read(ddir / "run.ini")
            ev = (ddir / "events.log").read_text().splitlines()
            last = int(ev[-1].split(": ", 1)[0])
            status = "completed"
            for line in ev:
                if "Simulation aborted" in line:
                    status = ("observation_aborted"
                              if any("Phase complete" in e for e in ev) else
                              "equilibration_aborted")
            rel = ddir.relative_to(sw)
            out.append(f"{rel},{cp['system']['temperature']},{cp['system']['density']},"
                       f"{cp['system']['random_seed']},\"{cp['equilibration']['name']}\","
                       f"\"{cp['observation']['name']}\",{last},{status}\n")
    (sdst / "runs.csv").write_text("".join(out))


if __name__ == "__main__":
    main(*sys.argv[1:])
