"""Writes the JSON fixtures of the reference's own thermochemistry unit tests (data only):
the gas and mechanism files those tests load.  Run where the reference tree is mounted;
the outputs are committed next to this script.

  O-thermo.json     src/gas/sample-data/O-thermo.lua          (cea_thermo_curves.d:186-217)
  O2-N2-H2.json     src/gas/sample-data/O2-N2-H2.lua          (therm_perf_gas_mix_eos.d:177-200)
  H2-I2-HI.json     src/kinetics/sample-input/H2-I2-HI.lua    (chemistry_update.d:658-751)
  H2-I2-inp.json    src/kinetics/sample-input/H2-I2-inp.lua   (chemistry_update.d:658-751)
  N2-diss.json      src/kinetics/sample-input/N2-diss.lua     (rate_constant.d:599-618)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from gdtk_amd.data.lua_tables import parse_file  # noqa: E402
from gdtk_amd.data.make_data import gas_from_db, thermo_table  # noqa: E402

REF = os.environ.get("GDTK_REFERENCE", "/root/reference")


def mech(path):
    env, _ = parse_file(path)
    sp = env["species"]
    species = [sp[k] for k in sorted(sp)] if isinstance(sp, dict) else sp
    out = []
    for k in sorted(env["reaction"]):
        r = env["reaction"][k]
        def side(idx, co):
            return [[int(i), int(round(c))] for i, c in zip(idx, co)]
        rr = {"equation": r["equation"], "type": r["type"],
              "reactants": side(r["reacIdx"], r["reacCoeffs"]),
              "products": side(r["prodIdx"], r["prodCoeffs"])}
        for key in ("frc", "brc"):
            rc = r[key]
            rr[key] = {"model": rc["model"]} if rc["model"] == "fromEqConst" else \
                {"model": rc["model"], "A": float(rc["A"]), "n": float(rc["n"]), "C": float(rc["C"])}
        if r["type"] == "anonymous_collider":
            rr["efficiencies"] = [[int(i), float(v)] for i, v in sorted(r["efficiencies"].items())]
        out.append(rr)
    cfg = env["config"]
    return {"species": species, "config": cfg, "reactions": out}


def main():
    env, _ = parse_file(os.path.join(REF, "src/gas/sample-data/O-thermo.lua"))
    json.dump({"R": 8.31451 / 0.0159994, "thermo": thermo_table(env["CEA_coeffs"])},
              open(os.path.join(HERE, "O-thermo.json"), "w"), indent=1)
    env, _ = parse_file(os.path.join(REF, "src/gas/sample-data/O2-N2-H2.lua"))
    sp = env["species"]
    json.dump({"species": sp, "M": [env[s]["M"] for s in sp], "thermo": [thermo_table(env[s]["cea_thermo"]) for s in sp]},
              open(os.path.join(HERE, "O2-N2-H2.json"), "w"), indent=1)
    json.dump(gas_from_db(os.path.join(REF, "src/kinetics/sample-input/H2-I2-HI.lua")),
              open(os.path.join(HERE, "H2-I2-HI.json"), "w"), indent=1)
    json.dump(mech(os.path.join(REF, "src/kinetics/sample-input/H2-I2-inp.lua")),
              open(os.path.join(HERE, "H2-I2-inp.json"), "w"), indent=1)
    env, _ = parse_file(os.path.join(REF, "src/kinetics/sample-input/N2-diss.lua"))
    json.dump(env["rate"], open(os.path.join(HERE, "N2-diss.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
