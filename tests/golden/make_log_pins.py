"""Extract V'(phi, T) values printed by the reference's simulation logs into a fixture.

Reads /root/reference/logs/*.log (text written by the reference drivers' start-up:
simulation/latticeSimeRescale_numba.py "Finding thermal VEV" / "Finding escape point" blocks and
the counterterm block at :758-800) and writes log_vprime_pins.json beside this script: per log
the temperature, couplings, tree VEV, potential type (null when the log predates the flag),
counterterm cutoff (dx_phys, T0) when applied with its printed soft-mode fractions, the printed
V'(phi) values (4-5 significant figures) and the barrier top.

Only logs that print the couplings and the tree VEV are kept.  Run in a container that has
the reference checkout; the fixture is what the tests read.
"""
import glob
import json
import os
import re
import sys

SRC = "/root/reference/logs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "log_vprime_pins.json")

NUM = r"([-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)"


def parse(path):
    txt = open(path, errors="replace").read()
    m_c = re.search(r"Couplings: boson=" + NUM + r", fermion=" + NUM, txt)
    m_v = re.search(r"Tree-level VEV: \S*?" + NUM, txt)
    m_t = re.search(r"Finding thermal VEV at T = " + NUM, txt)
    if not (m_c and m_v and m_t):
        return None
    rec = {"log": os.path.basename(path), "T": float(m_t.group(1)),
           "yb": float(m_c.group(1)), "yf": float(m_c.group(2)),
           "vev_tree": float(m_v.group(1)), "potential_type": None, "coeffs": None,
           "counterterm": None, "vprime": {}, "barrier_top": None}
    m_p = re.search(r"Potential type: (\S+)\s+\(coeffs: boson=" + NUM + r", fermion=" + NUM, txt)
    if m_p:
        rec["potential_type"] = m_p.group(1)
        rec["coeffs"] = [float(m_p.group(2)), float(m_p.group(3))]
    m_ct = re.search(r"dx_phys = " + NUM + r", T0 = " + NUM, txt)
    m_fr = re.search(r"Soft-mode fraction: boson=" + NUM + r" .*?fermion=" + NUM, txt)
    if m_ct and "Applying lattice counterterm" in txt:
        rec["counterterm"] = {"dx_phys": float(m_ct.group(1)), "T0": float(m_ct.group(2)),
                              "frac_b": float(m_fr.group(1)) if m_fr else None,
                              "frac_f": float(m_fr.group(2)) if m_fr else None}
    blk = txt[m_t.start():]
    end = blk.find("Using ")
    blk = blk[:end if end > 0 else 4000]
    for m in re.finditer(r"V'\((1e-?\d+)\)\s*=\s*" + NUM, blk):
        rec["vprime"][m.group(1)] = float(m.group(2))
    m_b = re.search(r"Barrier top at: " + NUM, blk)
    if m_b:
        rec["barrier_top"] = float(m_b.group(1))
    return rec if rec["vprime"] else None


def main():
    recs = [r for r in (parse(p) for p in sorted(glob.glob(os.path.join(SRC, "*.log")))) if r]
    # one record per distinct (T, couplings, type, counterterm) case
    seen, uniq = set(), []
    for r in recs:
        key = json.dumps([r["T"], r["yb"], r["yf"], r["vev_tree"], r["potential_type"],
                          r["counterterm"] and [r["counterterm"]["dx_phys"], r["counterterm"]["T0"]]])
        if key not in seen:
            seen.add(key)
            uniq.append(r)
    with open(OUT, "w") as fh:
        json.dump(uniq, fh, indent=1)
    print(f"{len(recs)} logs parsed, {len(uniq)} distinct cases -> {OUT}", file=sys.stderr)


if __name__ == "__main__":
    main()
