"""ORACLE — test infrastructure only (see oracle/__init__.py).

Python restatement of the Cantera-YAML front end: yaml.rs:40-171 (equation parsing, cm-mol-s unit
conversion A·(1e-6)^(order-1) with +1 order for three-body and the low-pressure limit, Ea cal/mol -> K,
active-first species order) and lib.rs:22-140 (molar masses from composition, NASA-7 ranges, SI transport
parameters, dense stoichiometry and efficiencies). Reads the file with PyYAML's SafeLoader.
"""
import math

import numpy as np
import yaml

KB = 1.380649e-23
NA = 6.02214076e23
R = KB * NA
ATOMIC_WEIGHTS = {"H": 1.008, "He": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998403163, "Cl": 35.45, "Ar": 39.95}
ELEMENTS = ["H", "He", "C", "N", "O", "F", "Cl", "Ar"]
MODELS = {"elementary": 0, "irreversible": 1, "three-body": 2, "pressure-modification": 3, "troe": 4, "sri": 5}


def _equation(text):
    sides = text.split("<=>") if "<=>" in text else text.split("=>")
    assert len(sides) == 2, text
    out = []
    for side in sides:
        side = side.strip()
        for suffix in ("(+ M)", "+ M", "(+M)"):
            while side.endswith(suffix):
                side = side[: -len(suffix)]
        d = {}
        for term in side.split(" + "):
            term = term.strip()
            if " " in term:
                count, specie = term.split()
                count = int(count)
            else:
                count, specie = 1, term
            d[specie] = d.get(specie, 0) + count
        out.append(d)
    return out


def _pressure(p):
    """A YAML pressure: a number (Pa) or '<number> <unit>' with unit Pa, kPa, MPa, bar, atm or torr."""
    if isinstance(p, (int, float)):
        return float(p)
    value, unit = (str(p).split() + ["Pa"])[:2]
    return float(value) * {"Pa": 1.0, "kPa": 1e3, "MPa": 1e6, "bar": 1e5, "atm": 101325.0, "torr": 101325.0 / 760.0}[unit]


def load(path=None, text=None):
    """Flat arrays of a mechanism (SI units), species in the reference's active-first order."""
    doc = yaml.safe_load(text if text is not None else open(path).read())
    units = doc["units"]
    assert units["length"] == "cm" and units["time"] == "s" and units["quantity"] == "mol"
    ea_unit = units["activation-energy"]
    species = []
    for s in doc["species"]:
        th = s["thermo"]
        ranges = [float(x) for x in th["temperature-ranges"]]
        pieces = [[float(x) for x in p] for p in th["data"]]
        if len(pieces) == 2 and pieces[0][:6] == pieces[1][:6] and abs(pieces[0][6] - pieces[1][6]) < 3e-8:
            ranges, pieces = [ranges[0], ranges[2]], [pieces[0]]
        species.append((s["name"], s, ranges, pieces))
    reactions = []
    for x in doc["reactions"]:
        eq = _equation(x["equation"])
        kind = x.get("type")
        order = sum(eq[0].values()) + (1 if kind == "three-body" else 0)

        def rate(k, exponent):
            Ea = float(k["Ea"])
            return [float(k["A"]) * math.pow(1e-6, exponent), float(k["b"]), Ea if ea_unit == "K" else Ea * 4.184 / R]

        rc = x.get("rate-constant") or x.get("high-P-rate-constant") or x["rate-constants"][0]
        plog = []
        if kind == "pressure-dependent-Arrhenius" and len(x["rate-constants"]) > 1:
            plog = sorted(([_pressure(e["P"])] + rate(e, order - 1) for e in x["rate-constants"]), key=lambda row: row[0])
        entry = {"eq": eq, "rate": rate(rc, order - 1), "k0": [0.0, 0.0, 0.0], "troe": [0.0, 0.0, 0.0, 0.0], "sri": [0.0] * 5, "plog": plog,
                 "eff": {k: float(v) for k, v in (x.get("efficiencies") or {}).items()}}
        reversible = " <=> " in x["equation"]
        if kind in (None, "elementary", "pressure-dependent-Arrhenius"):
            entry["model"] = "elementary" if reversible else "irreversible"
        elif kind == "three-body":
            entry["model"] = "three-body"
        elif kind == "falloff":
            entry["k0"] = rate(x["low-P-rate-constant"], order)
            if "Troe" in x:
                t = x["Troe"]
                entry["model"] = "troe"
                entry["troe"] = [float(t["A"]), float(t["T3"]), float(t["T1"]), float(t.get("T2", math.inf))]
            elif "SRI" in x:
                s = x["SRI"]
                entry["model"] = "sri"
                entry["sri"] = [float(s["A"]), float(s["B"]), float(s["C"]), float(s.get("D", 1.0)), float(s.get("E", 0.0))]
            else:
                entry["model"] = "pressure-modification"
        else:
            raise ValueError(kind)
        reactions.append(entry)

    def reacts(name):
        return any(e["eq"][0].get(name, 0) != e["eq"][1].get(name, 0) for e in reactions)

    species = [s for s in species if reacts(s[0])] + [s for s in species if not reacts(s[0])]
    names = [s[0] for s in species]
    K = len(names)
    Ka = sum(reacts(n) for n in names)
    # lib.rs:23,42: each atomic weight in kg/mol, summed in composition order
    W = np.array([sum(n * (ATOMIC_WEIGHTS[e] / 1e3) for e, n in s["composition"].items()) for _, s, _, _ in species])
    composition = np.array([[s["composition"].get(e, 0) for e in ELEMENTS] for _, s, _, _ in species])
    split = np.array([r[1] if len(r) == 3 else math.nan for _, _, r, _ in species])
    nasa = np.array([[p[0], p[1] if len(p) == 2 else p[0]] for _, _, _, p in species])
    geometry = [s["transport"]["geometry"] for _, s, _, _ in species]
    tr = [s["transport"] for _, s, _, _ in species]
    Rn = len(reactions)
    reactants = np.zeros((Rn, K), np.int32)
    products = np.zeros((Rn, K), np.int32)
    eff = np.ones((Rn, K))
    for r, e in enumerate(reactions):
        for i, name in enumerate(names):
            reactants[r, i] = e["eq"][0].get(name, 0)
            products[r, i] = e["eq"][1].get(name, 0)
            if e["model"] in ("three-body", "pressure-modification", "troe", "sri"):
                eff[r, i] = e["eff"].get(name, 1.0)
    return {
        "names": names, "K": K, "Ka": Ka, "R": Rn, "W": W, "composition": composition, "split": split, "nasa": nasa,
        "reactants": reactants, "products": products,
        "model": np.array([MODELS[e["model"]] for e in reactions], np.int32),
        "rate": np.array([e["rate"] for e in reactions]), "k0": np.array([e["k0"] for e in reactions]),
        "troe": np.array([e["troe"] for e in reactions]), "sri": np.array([e["sri"] for e in reactions]), "eff": eff,
        "plog": [e["plog"] for e in reactions],
        "geometry": geometry,
        "diameter": np.array([t["diameter"] * 1e-10 for t in tr]),
        "well_depth_J": np.array([t["well-depth"] * KB for t in tr]),
        "polarizability": np.array([0.0 if g == "atom" else t.get("polarizability", 0.0) * 1e-30 for t, g in zip(tr, geometry)]),
        "dipole": np.array([t.get("dipole", 0.0) * 1e-21 / 299792458.0 if g == "nonlinear" else 0.0 for t, g in zip(tr, geometry)]),
        "rotational_relaxation": np.array([0.0 if g == "atom" else t.get("rotational-relaxation", 0.0) for t, g in zip(tr, geometry)]),
        "internal_dof": np.array([{"atom": 0.0, "linear": 1.0, "nonlinear": 1.5}[g] for g in geometry]),
    }
