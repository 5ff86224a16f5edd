"""Writes tests/golden/bbB76_wet_trace.json: the wet rows (box 76, before the first drying ramp) of
the reference's box-76 campaign trace, verify/out/big-box-campaign-B76-trace.json -- the campaign
`--box 76 --start {C 55988, O 18659, H 55988, M 1452, W 351181} --seed 19 --kT 1.1 --relax --cycle
--evaporate --cycles 1 --charge --pH 7.0 --ionicStrength 0.01` (docs/verdict.md:413-416). Run in a
container that holds the reference; the fixture is data only (step, events, bonds, amphiphiles)."""
import json
import os
import sys

SRC = "/root/reference/verify/out/big-box-campaign-B76-trace.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bbB76_wet_trace.json")


def main():
    rows = json.load(open(SRC))
    keep = [{"step": r["step"], "N": r["N"], "events": r["events"], "bonds": r["bonds"],
             "amphiphileCount": r["amphiphileCount"], "stage": r["stage"],
             "valence": r["invariants"]["valence"]} for r in rows if r["box"] == [76, 76, 76] and r["step"] < 20000]
    doc = {"source": "verify/out/big-box-campaign-B76-trace.json", "campaign": "bbB76 (docs/verdict.md:142-145, :413-416)",
           "rows": keep}
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
        fh.write("\n")
    print(f"wrote {len(keep)} rows to {OUT}", file=sys.stderr)


if __name__ == "__main__":
    main()
