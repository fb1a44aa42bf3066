"""Writes the JSON gas-model and mechanism data of the chemistry path from the reference's
Lua data files (run once where the reference tree is mounted; the JSON is committed).

  tpg-air-5sp.json      thermally perfect 5-species air, species N2, O2, NO, N, O
                        (src/gas/sample-data/therm-perf-5-species-air.lua)
  tpg-nitrogen-2sp.json N2, N of the same curves (the n90 cylinder's nitrogen-2sp.inp)
  nitrogen-2sp-2r.json  the n90 cylinder's 2-reaction N2 dissociation mechanism
  park-air-5sp-1T.json  one-temperature Park 5-reaction mechanism authored from
                        examples/lmr/3D/cylinder-larc/rm-park-5sp-2T.inp: the 'Park'
                        rate constants (T^s Tv^(1-s), identical to T when Tv = T) become
                        'Arrhenius' with the same A, n, C; A is converted from CGS to SI
                        with prep-chem's factor 1e-6^(sum of reactant coefficients + 1 for
                        an anonymous collider - 1) (src/kinetics/reaction.lua:26-45);
                        backward rates come from the equilibrium constant; the ODE
                        settings are prep-chem's defaults for alpha-QSS
                        (src/kinetics/prep_chem.lua:46-90).
  gupta-air-5sp-2T.json the 5-species form of examples/eilmer/2D/fireII/legacy/
                        air-11sp-reactions.inp (NO_SPECIES = 5 selects r1-r6): Park rates
                        at T^s Tv^(1-s) forward, Arrhenius backward, converted to SI as
                        prep-chem does (A written "%.12e", n and s "%f"); efficiencies
                        default to 1, the file's values override, zeros are dropped, as
                        prep-chem writes them (src/kinetics/reaction.lua:514-545, 583-596);
                        the file's Config (tempLimits 300-30000 K, RKF errTol 1e-9, tight
                        temperature coupling) with prep-chem's default step limits.
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from gdtk_amd.data.lua_tables import parse, parse_file  # noqa: E402

REF = os.environ.get("GDTK_REFERENCE", "/root/reference")


def thermo_table(tc):
    nseg = tc["nsegments"]
    return {"T_breaks": [float(x) for x in tc["T_break_points"]],
            "T_blends": [float(x) for x in tc["T_blend_ranges"]],
            "coeffs": [[float(x) for x in tc[f"segment{i}"]] for i in range(nseg)]}


def gas_from_db(path, thermo_key="thermoCoeffs"):
    env, _ = parse_file(path)
    species = env["species"]
    if isinstance(species, dict):
        species = [species[k] for k in sorted(species)]
    db = env["db"] if "db" in env else env
    return {"model": "ThermallyPerfectGas", "source": os.path.relpath(path, REF), "species": species,
            "M": [float(db[s]["M"]) for s in species],
            "thermo": [thermo_table(db[s][thermo_key]) for s in species]}


def parse_side(side, species):
    """'N + N + M' -> ({isp: coeff}, has_collider)"""
    coeffs, collider = {}, False
    for term in side.split("+"):
        term = term.strip()
        m = re.fullmatch(r"(\d+)?\s*([A-Za-z0-9_]+)", term)
        n, name = int(m.group(1) or 1), m.group(2)
        if name == "M":
            collider = True
            continue
        isp = species.index(name)
        coeffs[isp] = coeffs.get(isp, 0) + n
    return coeffs, collider


def park_1T(gas):
    species = gas["species"]
    _, calls = parse_file(os.path.join(REF, "examples/lmr/3D/cylinder-larc/rm-park-5sp-2T.inp"))
    reactions = []
    for name, t in calls:
        assert name == "Reaction"
        eq = t[1]
        lhs, rhs = eq.split("<=>")
        reac, colL = parse_side(lhs, species)
        prod, colR = parse_side(rhs, species)
        assert colL == colR
        fr = t["fr"]
        nc = sum(reac.values()) + (1 if colL else 0)
        conv = 1.0e-6 ** (nc - 1)
        r = {"equation": eq, "label": t.get("label", ""),
             "type": "anonymous_collider" if colL else "elementary",
             "reactants": [[i, c] for i, c in sorted(reac.items())],
             "products": [[i, c] for i, c in sorted(prod.items())],
             # prep-chem writes the converted A with "%.12e" (reaction.lua, transformRateConstant).
             "frc": {"model": "Arrhenius", "A": float("%.12e" % (fr["A"] * conv)), "n": float(fr["n"]),
                     "C": float(fr["C"]),
                     "A_cgs": fr["A"], "source_model": fr[1]},
             "brc": {"model": "fromEqConst"}}
        if colL:
            eff = t.get("efficiencies", {})
            r["efficiencies"] = [[species.index(s), float(v)] for s, v in
                                 sorted(eff.items(), key=lambda kv: species.index(kv[0]))]
        reactions.append(r)
    return {"species": species, "source": "examples/lmr/3D/cylinder-larc/rm-park-5sp-2T.inp (1-T Arrhenius form)",
            "config": {"tempLimits": {"lower": 300.0, "upper": 50000.0},
                       "odeStep": {"method": "alpha-qss", "eps1": 1.0e-3, "eps2": 5.0e-4, "delta": 1.0e-10,
                                   "maxIters": 10},
                       "tightTempCoupling": False, "maxSubcycles": 10000, "maxAttempts": 4},
            "reactions": reactions}


def arrhenius_si(t, ncoeff):
    """An 'Arrhenius' rate of a Lua Reaction table in SI units (prep-chem's conversion,
    src/kinetics/reaction.lua:26-45, 486-491, written with "%.12e")."""
    conv = 1.0e-6 ** (ncoeff - 1)
    return {"model": "Arrhenius", "A": float("%.12e" % (t["A"] * conv)), "n": float(t["n"]), "C": float(t["C"]),
            "A_cgs": t["A"]}


def nitrogen_2sp_2r(species):
    """examples/eilmer/2D/cylinder-dlr-n90/cpu-chem/nitrogen-2sp-2r.lua: N2 dissociation by
    N2 and by N, both directions given as Arrhenius rates (elementary reactions, no
    anonymous collider)."""
    _, calls = parse_file(os.path.join(REF, "examples/eilmer/2D/cylinder-dlr-n90/cpu-chem/nitrogen-2sp-2r.lua"))
    reactions = []
    for name, t in calls:
        assert name == "Reaction"
        eq = t[1]
        lhs, rhs = eq.split("<=>")
        reac, colL = parse_side(lhs, species)
        prod, colR = parse_side(rhs, species)
        assert not colL and not colR
        reactions.append({"equation": eq, "label": "", "type": "elementary",
                          "reactants": [[i, c] for i, c in sorted(reac.items())],
                          "products": [[i, c] for i, c in sorted(prod.items())],
                          "frc": arrhenius_si(t["fr"], sum(reac.values())),
                          "brc": arrhenius_si(t["br"], sum(prod.values()))})
    return {"species": species, "source": "examples/eilmer/2D/cylinder-dlr-n90/cpu-chem/nitrogen-2sp-2r.lua",
            "config": {"tempLimits": {"lower": 300.0, "upper": 50000.0},
                       "odeStep": {"method": "alpha-qss", "eps1": 1.0e-3, "eps2": 5.0e-4, "delta": 1.0e-10,
                                   "maxIters": 10},
                       "tightTempCoupling": False, "maxSubcycles": 10000, "maxAttempts": 4},
            "reactions": reactions}


def _reaction_blocks(text):
    """The top-level ``Reaction{...}`` statements of a mechanism input, by brace matching
    (the rest of such a file -- variables, ``if`` blocks -- is not evaluated)."""
    out, pos = [], 0
    while True:
        k = text.find("Reaction{", pos)
        if k < 0:
            return out
        depth, j = 0, k + len("Reaction")
        while True:
            c = text[j]
            if c == "{":
                depth += 1
            elif c == "}":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        out.append(text[k:j + 1])
        pos = j + 1


def _fmt_f(x):
    return float("%f" % x)


def gupta_2T(gas):
    species = gas["species"]
    path = os.path.join(REF, "examples/eilmer/2D/fireII/legacy/air-11sp-reactions.inp")
    with open(path) as f:
        text = f.read()
    s_val = float(re.search(r"^s\s*=\s*([0-9.eE+-]+)", text, re.M).group(1))
    reactions = []
    for block in _reaction_blocks(text):
        _, calls = parse(re.sub(r"\bs=s\b", f"s={s_val!r}", block))
        (name, t), = calls
        if t.get("label") not in ("r1", "r2", "r3", "r4", "r5", "r6"):
            continue
        eq = t[1]
        lhs, rhs = eq.split("<=>")
        reac, colL = parse_side(lhs, species)
        prod, colR = parse_side(rhs, species)
        assert colL == colR
        nf = sum(reac.values()) + (1 if colL else 0)
        nb = sum(prod.values()) + (1 if colR else 0)

        def rate(rt, nc):
            conv = 1.0e-6 ** (nc - 1)
            d = {"model": rt[1], "A": float("%.12e" % (rt["A"] * conv)), "n": _fmt_f(rt["n"]),
                 "C": float("%.12e" % rt["C"]), "A_cgs": rt["A"]}
            if rt[1] == "Park":
                d["s"] = _fmt_f(rt["s"])
                d["mode"] = 0
            else:
                assert rt[1] == "Arrhenius"
            return d

        r = {"equation": eq, "label": t["label"], "type": "anonymous_collider" if colL else "elementary",
             "reactants": [[i, c] for i, c in sorted(reac.items())],
             "products": [[i, c] for i, c in sorted(prod.items())],
             "frc": rate(t["fr"], nf), "brc": rate(t["br"], nb)}
        if colL:
            eff = {i: 1.0 for i in range(len(species))}
            for sp, v in t.get("efficiencies", {}).items():
                if sp in species:
                    eff[species.index(sp)] = float(v)
            r["efficiencies"] = [[i, float("%.6e" % v)] for i, v in sorted(eff.items()) if v != 0.0]
        reactions.append(r)
    assert [r["label"] for r in reactions] == ["r1", "r2", "r3", "r4", "r5", "r6"]
    return {"species": species, "source": "examples/eilmer/2D/fireII/legacy/air-11sp-reactions.inp (NO_SPECIES = 5)",
            "config": {"tempLimits": {"lower": 300.0, "upper": 30000.0},
                       "odeStep": {"method": "rkf", "errTol": 1.0e-9},
                       "tightTempCoupling": True, "maxSubcycles": 10000, "maxAttempts": 4},
            "two_temperature": {"molecular": [s in ("N2", "O2", "NO") for s in species],
                                "A_MW": [{"N2": 220.0, "O2": 129.0, "NO": 168.0}.get(s, 0.0) for s in species],
                                "source": "src/kinetics/two_temperature_air_kinetics.d:714-720 (Gnoffo et al. "
                                          "Table 1), src/gas/two_temperature_air.d:25"},
            "reactions": reactions}


def subset_gas(gas, species, source):
    idx = [gas["species"].index(s) for s in species]
    return {"model": gas["model"], "source": source, "species": list(species), "M": [gas["M"][i] for i in idx],
            "thermo": [gas["thermo"][i] for i in idx]}


def main():
    gas = gas_from_db(os.path.join(REF, "src/gas/sample-data/therm-perf-5-species-air.lua"))
    with open(os.path.join(HERE, "tpg-air-5sp.json"), "w") as f:
        json.dump(gas, f, indent=1)
    with open(os.path.join(HERE, "park-air-5sp-1T.json"), "w") as f:
        json.dump(park_1T(gas), f, indent=1)
    # nitrogen-2sp.inp (N2, N) takes the species database's CEA curves, which are the ones
    # of the 5-species sample gas (same database, same prep-gas output).
    n2 = subset_gas(gas, ["N2", "N"], "examples/eilmer/2D/cylinder-dlr-n90/cpu-chem/nitrogen-2sp.inp "
                                      "(CEA curves as in src/gas/sample-data/therm-perf-5-species-air.lua)")
    with open(os.path.join(HERE, "tpg-nitrogen-2sp.json"), "w") as f:
        json.dump(n2, f, indent=1)
    with open(os.path.join(HERE, "nitrogen-2sp-2r.json"), "w") as f:
        json.dump(nitrogen_2sp_2r(n2["species"]), f, indent=1)
    with open(os.path.join(HERE, "gupta-air-5sp-2T.json"), "w") as f:
        json.dump(gupta_2T(gas), f, indent=1)


if __name__ == "__main__":
    main()
