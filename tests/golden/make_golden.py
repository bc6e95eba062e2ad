"""Collect the reference's own regression values for the chemistry path into
small fixtures (run once in a tree that has the reference checkout):

    python tests/golden/make_golden.py /root/reference

Writes reference_golds.json (values copied from the *.regression.gold files)
and copies the thermodynamic databases the decks use (data files held by the
reference's tests).  The tests read only the fixtures.
"""
import json
import os
import shutil
import sys

CASES = {
    # deck -> (gold file, database)
    "carbonate-debye-huckel-activity": ("regression_tests/ascem/batch/"
                                        "carbonate-debye-huckel-activity.regression.gold",
                                        "regression_tests/ascem/batch/carbonate.dat"),
    "carbonate-unit-activity": ("regression_tests/ascem/batch/"
                                "carbonate-unit-activity.regression.gold",
                                "regression_tests/ascem/batch/carbonate.dat"),
    "calcite-kinetics": ("regression_tests/ascem/batch/calcite-kinetics.regression.gold",
                         "regression_tests/ascem/batch/calcite.dat"),
    # initial state of a three-component deck: the constraint speciation's output
    "ca-carbonate-debye-huckel-activity": ("regression_tests/ascem/batch/"
                                           "ca-carbonate-debye-huckel-activity.regression.gold",
                                           "regression_tests/ascem/batch/ca-carbonate.dat"),
    "ca-carbonate-unit-activity": ("regression_tests/ascem/batch/"
                                   "ca-carbonate-unit-activity.regression.gold",
                                   "regression_tests/ascem/batch/ca-carbonate.dat"),
}


def parse_gold(path):
    """{'GENERIC: pH': value of cell 1, ...} plus the solver summary counts"""
    out = {}
    key = None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("--") and s.endswith("--"):
                key = s[2:-2].strip()
                continue
            if key is None or ":" not in s:
                continue
            k, v = [t.strip() for t in s.split(":", 1)]
            if k == "1":
                out[key] = float(v)
            elif key.startswith("SOLUTION") and k in ("Time Steps", "Newton Iterations",
                                                      "Time Step Cuts"):
                out["%s: %s" % (key, k)] = float(v)
    return out


def parse_obs(path, nrows):
    """header and the first nrows rows of a *-obs-0.tec.gold time series"""
    with open(path) as f:
        header = [h.strip().strip('"') for h in f.readline().split('",')]
        rows = []
        for line in f:
            if len(rows) == nrows:
                break
            rows.append([float(v) for v in line.split()])
    return header, rows


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    golds = {}
    # SomDec cascade deck: the first steps of the observation time series
    # (7 significant digits) and the end-of-run regression values
    deck = "regression_tests/ngee/CLMCNplus/clm_cn3"
    header, rows = parse_obs(os.path.join(ref, deck + "-obs-0.tec.gold"), 25)
    golds["clm_cn3"] = {"source": deck + "-obs-0.tec.gold",
                        "columns": header, "rows": rows,
                        "final": parse_gold(os.path.join(ref, deck + ".regression.gold"))}
    # one-reaction SOM decks with N2O emission: end-of-run regression values;
    # clm_som1 also the step times of its observation series (every step)
    deck = "regression_tests/ngee/CLMCNplus/clm_som1"
    header, rows = parse_obs(os.path.join(ref, deck + "-obs-0.tec.gold"), 10 ** 6)
    golds["clm_som1"] = {"source": deck + ".regression.gold",
                         "obs_times_d": [r[0] for r in rows],
                         "final": parse_gold(os.path.join(ref, deck + ".regression.gold"))}
    deck = "regression_tests/ngee/CLMCNplus/clm_nmin"
    golds["clm_nmin"] = {"source": deck + ".regression.gold",
                         "final": parse_gold(os.path.join(ref, deck + ".regression.gold"))}
    for case, (gold, db) in CASES.items():
        golds[case] = {"source": gold, "database": os.path.basename(db),
                       "values": parse_gold(os.path.join(ref, gold))}
        shutil.copyfile(os.path.join(ref, db), os.path.join(here, os.path.basename(db)))
    with open(os.path.join(here, "reference_golds.json"), "w") as f:
        json.dump(golds, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
