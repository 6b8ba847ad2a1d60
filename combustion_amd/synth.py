"""Seeded synthetic mechanisms and states, written as Cantera YAML and SoA arrays.

The benchmark mechanisms are drawn, not read: no mechanism file ships with the reference (its mechanisms
come from a Cantera installation, cantera.rs:7,15). Each draw is fixed by its seed. Choices the parameter
ranges leave open are frozen here:

* Species are H/C/O molecules (small named ones first, then random formulas), plus an inert bulk N2 last.
  Reactions are element balanced (so Σ_k W_k ω_k = 0 holds exactly in exact arithmetic) with 1-3 species per
  side, |ν| <= 2 and a change in moles Σν in {-1, 0, 1, 2} (reaction.rs:298-303 supports -Σν in {0, ±1, -2}).
  Every reactive species changes in some reaction, so the reacting species form a prefix (lib.rs:130-136).
* Arrhenius: ln A ~ U[20, 40] in cm-mol-s units of the reaction order, β ~ U[-1.5, 2.5], Ea/R ~ U[0, 25000] K.
  Three-body and falloff reactions are recombinations A + B (+M) <=> AB; each has 6 collision efficiencies
  ~ U[0.5, 12] on random species (the bulk included). Troe: α ~ U[0.1, 0.9], T3 ~ U[50, 1000] K,
  T1 ~ U[500, 5000] K, T2 ~ U[1000, 8000] K on half of them (absent on the others); k0 has its own draw.
* NASA-7: a smooth increasing cp/R(T) = a + (b - a)·σ((T - Tm)/w) with 3 <= a < b <= 8, fitted by least squares
  on 300-1000 K and 1000-3500 K; h_f/R(298.15 K) ~ U[-3e4, 3e4] K and s°/R(298.15 K) ~ U[20, 40] fix a5, a6 of
  the low piece; the high piece's a5, a6 make h and s continuous at 1000 K.
* Transport: σ ~ U[2.5, 4.5] Å, ε/k ~ U[30, 600] K, no dipole (δ* = 0), polarizability ~ U[0, 3] Å³,
  rotational relaxation ~ U[0, 5]; geometry atom / linear (2 atoms) / nonlinear. The well depth starts at
  30 K, not 10 K: the fits sample 300-3000 K (transport.rs:127) and the Ω*22 table ends at T* = kT/ε = 100
  (transport/collision_integrals.rs:3-41), so a pair with ε/k < 30 K leaves the table, which the reference
  rejects (transport.rs:87).
"""
from __future__ import annotations

import itertools
import math
import re
import os

import numpy as np

NAMED = ["H", "O", "H2", "O2", "OH", "H2O", "HO2", "H2O2", "O3", "CO", "CO2", "HCO", "CH2O", "CH4", "CH3", "CH2", "CH",
         "CH3O", "CH2OH", "CH3OH", "C2H2", "C2H3", "C2H4", "C2H5", "C2H6", "HCCO", "CH2CO", "HCCOH", "C2H", "C2H4O",
         "C2H5O", "C3H3", "C3H4", "C3H5", "C3H6", "C3H7", "C3H8", "C2O", "HO3", "H2O3"]
ATOMIC_WEIGHT = {"H": 1.008, "C": 12.011, "O": 15.999, "N": 14.007}


def parse_formula(name):
    out, i = {}, 0
    while i < len(name):
        e = name[i]
        i += 1
        j = i
        while j < len(name) and name[j].isdigit():
            j += 1
        n = int(name[i:j]) if j > i else 1
        out[e] = out.get(e, 0) + n
        i = j
    return out


def formula(comp):
    s = ""
    for e in ("C", "H", "O"):
        n = comp.get(e, 0)
        if n:
            s += e + (str(n) if n > 1 else "")
    return s


class Mechanism:
    """A drawn mechanism: species compositions, reactions and all parameters, plus its Cantera YAML."""

    def __init__(self, species, reactions, meta):
        self.species = species      # list of dicts (name, composition, nasa, transport)
        self.reactions = reactions  # list of dicts
        self.meta = meta

    @property
    def K(self):
        return len(self.species)

    def yaml(self):
        return to_yaml(self)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _species_compositions(n_reactive, rng):
    comps = [parse_formula(n) for n in NAMED[:n_reactive]]
    names = list(NAMED[:n_reactive])
    seen = {}
    for n, c in zip(names, comps):
        seen[tuple(sorted(c.items()))] = seen.get(tuple(sorted(c.items())), 0) + 1
    while len(comps) < n_reactive:
        c = {"C": int(rng.integers(0, 4)), "H": int(rng.integers(0, 9)), "O": int(rng.integers(0, 4))}
        c = {k: v for k, v in c.items() if v}
        if sum(c.values()) < 2 or sum(c.values()) > 10:
            continue
        key = tuple(sorted(c.items()))
        count = seen.get(key, 0)
        if count >= 3:
            continue
        seen[key] = count + 1
        name = formula(c) + ("" if count == 0 else "-" + "abc"[count - 1])
        if name in names:
            continue
        names.append(name)
        comps.append(c)
    return names, comps


def _nasa7(rng):
    a = rng.uniform(3.0, 4.5)
    b = rng.uniform(max(a + 0.5, 5.0), 8.0)
    Tm = rng.uniform(500.0, 1500.0)
    w = rng.uniform(200.0, 600.0)
    cp = lambda T: a + (b - a) / (1.0 + np.exp(-(T - Tm) / w))
    pieces = []
    for lo, hi in ((300.0, 1000.0), (1000.0, 3500.0)):
        T = np.linspace(lo, hi, 200)
        V = np.stack([T ** i for i in range(5)], axis=1)
        s = np.array([1.0, 1e-3, 1e-6, 1e-9, 1e-12])  # column scaling for conditioning
        coef = np.linalg.lstsq(V * s, cp(T), rcond=None)[0] * s
        pieces.append([float(x) for x in coef])

    def h_R(p, T):  # without a5
        return p[0] * T + p[1] * T ** 2 / 2 + p[2] * T ** 3 / 3 + p[3] * T ** 4 / 4 + p[4] * T ** 5 / 5

    def s_R(p, T):  # without a6
        return p[0] * math.log(T) + p[1] * T + p[2] * T ** 2 / 2 + p[3] * T ** 3 / 3 + p[4] * T ** 4 / 4

    hf = rng.uniform(-3e4, 3e4)
    s0 = rng.uniform(20.0, 40.0)
    lo, hi = pieces
    a5 = hf - h_R(lo, 298.15)
    a6 = s0 - s_R(lo, 298.15)
    lo = lo + [a5, a6]
    hi = hi + [h_R(lo, 1000.0) + a5 - h_R(hi, 1000.0), s_R(lo, 1000.0) + a6 - s_R(hi, 1000.0)]
    return [300.0, 1000.0, 3500.0], [lo, hi]


def _transport(comp, rng):
    atoms = sum(comp.values())
    t = {"well-depth": float(rng.uniform(30.0, 600.0)), "diameter": float(rng.uniform(2.5, 4.5))}
    if atoms == 1:
        t["geometry"] = "atom"
    else:
        t["geometry"] = "linear" if atoms == 2 else "nonlinear"
        t["polarizability"] = float(rng.uniform(0.0, 3.0))
        t["rotational-relaxation"] = float(rng.uniform(0.0, 5.0))
    return t


def _code(c):
    return c.get("C", 0) * 10000 + c.get("H", 0) * 100 + c.get("O", 0)


class _Balanced:
    """Samples element-balanced (reactants, products) multisets of 1-3 species."""

    def __init__(self, comps, rng):
        self.rng = rng
        self.n = len(comps)
        self.codes = [_code(c) for c in comps]
        self.groups = {1: {}, 2: {}, 3: {}}
        for size in (1, 2, 3):
            for ms in itertools.combinations_with_replacement(range(self.n), size):
                if size == 3 and ms[0] == ms[2]:
                    continue
                self.groups[size].setdefault(sum(self.codes[i] for i in ms), []).append(ms)

    def partner(self, r, size):
        """A product multiset of the given size, balanced with r and sharing no species with it."""
        items = self.groups[size].get(sum(self.codes[i] for i in r), [])
        if not items:
            return None
        for _ in range(8):
            p = items[int(self.rng.integers(len(items)))]
            if not set(p) & set(r):
                return p
        return None

    def draw(self, sizes, containing=None):
        """(reactants, products) with (len r, len p) drawn from sizes, optionally containing a species."""
        for _ in range(4000):
            lr, lp = sizes[int(self.rng.integers(len(sizes)))]
            extra = [int(x) for x in self.rng.integers(self.n, size=lr - (containing is not None))]
            r = tuple(sorted(extra + ([containing] if containing is not None else [])))
            if lr == 3 and r[0] == r[2]:
                continue
            p = self.partner(r, lp)
            if p is not None:
                return r, p
        return None


def generate(n_species, n_reactions, n_three_body, n_troe, seed, name="synthetic"):
    """Draws a mechanism of n_species (the last one inert N2) and n_reactions reactions."""
    rng = _rng(seed)
    n_reactive = n_species - 1
    names, comps = _species_compositions(n_reactive, rng)
    sampler = _Balanced(comps, rng)
    n_falloff = n_three_body + n_troe
    used = set()
    other_sizes = [(2, 2), (2, 2), (2, 2), (1, 2), (2, 3), (1, 1), (3, 2), (1, 3), (2, 1)]

    def fresh(pair):
        if pair is None:
            return False
        if pair in used or (pair[1], pair[0]) in used:
            return False
        used.add(pair)
        return True

    def draw(sizes, count, containing=None):
        out = []
        for _ in range(50 * count + 100):
            pair = sampler.draw(sizes, containing)
            if fresh(pair):
                out.append(pair)
                if len(out) == count:
                    return out
        raise ValueError("not enough balanced reactions for this species set")

    def changed(r, p):
        return {x for x in set(r) | set(p) if r.count(x) != p.count(x)}

    falloff = draw([(2, 1)], n_falloff) if n_falloff else []
    covered = set()
    for r, p in falloff:
        covered |= changed(r, p)
    elementary = []
    for s in rng.permutation(n_reactive):
        s = int(s)
        if s in covered:
            continue
        pair = draw(other_sizes, 1, containing=s)[0]
        elementary.append(pair)
        covered |= changed(*pair)
    n_elementary = n_reactions - n_falloff
    if len(elementary) > n_elementary:
        raise ValueError("too many reactions needed for coverage")
    if n_elementary > len(elementary):
        elementary += draw(other_sizes, n_elementary - len(elementary))
    species = []
    for nm, c in zip(names, comps):
        ranges, data = _nasa7(rng)
        species.append({"name": nm, "composition": c, "temperature-ranges": ranges, "data": data, "transport": _transport(c, rng)})
    ranges, data = _nasa7(rng)
    species.append({"name": "N2", "composition": {"N": 2}, "temperature-ranges": ranges, "data": data,
                    "transport": _transport({"N": 2}, rng)})
    all_names = [s["name"] for s in species]

    def arrhenius(order):
        return {"A": float(math.exp(rng.uniform(20.0, 40.0))), "b": float(rng.uniform(-1.5, 2.5)),
                "Ea": float(rng.uniform(0.0, 25000.0))}

    def efficiencies():
        idx = rng.choice(n_species, size=min(6, n_species), replace=False)
        return {all_names[i]: float(rng.uniform(0.5, 12.0)) for i in sorted(idx)}

    reactions = []
    kinds = ["three-body"] * n_three_body + ["troe"] * n_troe
    for (r, p), kind in zip(falloff, kinds):
        reactions.append({"reactants": r, "products": p, "type": kind})
    for r, p in elementary:
        reactions.append({"reactants": r, "products": p, "type": "elementary"})
    order = rng.permutation(len(reactions))
    reactions = [reactions[i] for i in order]
    for x in reactions:
        x["rate-constant"] = arrhenius(len(x["reactants"]))
        if x["type"] in ("three-body", "troe"):
            x["efficiencies"] = efficiencies()
        if x["type"] == "troe":
            x["low-P-rate-constant"] = arrhenius(len(x["reactants"]) + 1)
            t = {"A": float(rng.uniform(0.1, 0.9)), "T3": float(rng.uniform(50.0, 1000.0)), "T1": float(rng.uniform(500.0, 5000.0))}
            if rng.uniform() < 0.5:
                t["T2"] = float(rng.uniform(1000.0, 8000.0))
            x["Troe"] = t
        x["equation"] = _equation(x, names)
    return Mechanism(species, reactions, {"name": name, "seed": seed, "n_species": n_species, "n_reactions": n_reactions,
                                          "n_three_body": n_three_body, "n_troe": n_troe})


def _side(ms, names):
    counts = {}
    for i in ms:
        counts[i] = counts.get(i, 0) + 1
    return " + ".join((f"{c} {names[i]}" if c > 1 else names[i]) for i, c in counts.items())


def _equation(x, names):
    lhs, rhs = _side(x["reactants"], names), _side(x["products"], names)
    if x["type"] == "three-body":
        return f"{lhs} + M <=> {rhs} + M"
    if x["type"] == "troe":
        return f"{lhs} (+M) <=> {rhs} (+M)"
    return f"{lhs} <=> {rhs}"


def _num(x):
    return repr(float(x))


def _map(d):
    return "{" + ", ".join(f"{k}: {_num(v) if isinstance(v, float) else v}" for k, v in d.items()) + "}"


def to_yaml(m):
    lines = [f"description: synthetic mechanism {m.meta['name']} (seed {m.meta['seed']})",
             "units: {length: cm, time: s, quantity: mol, activation-energy: K}", "phases:", "- name: gas",
             "  thermo: ideal-gas", "  elements: [H, C, O, N]",
             "  species: [" + ", ".join(s["name"] for s in m.species) + "]", "  kinetics: gas",
             "  transport: mixture-averaged", "species:"]
    for s in m.species:
        lines += [f"- name: {s['name']}", f"  composition: {_map(s['composition'])}", "  thermo:", "    model: NASA7",
                  "    temperature-ranges: [" + ", ".join(_num(t) for t in s["temperature-ranges"]) + "]", "    data:"]
        for p in s["data"]:
            lines.append("    - [" + ", ".join(_num(a) for a in p) + "]")
        t = s["transport"]
        lines += ["  transport:", "    model: gas", f"    geometry: {t['geometry']}", f"    well-depth: {_num(t['well-depth'])}",
                  f"    diameter: {_num(t['diameter'])}"]
        for key in ("polarizability", "rotational-relaxation", "dipole"):
            if key in t:
                lines.append(f"    {key}: {_num(t[key])}")
    lines.append("reactions:")
    for x in m.reactions:
        lines.append(f"- equation: {x['equation']}")
        if x["type"] == "three-body":
            lines.append("  type: three-body")
        if x["type"] == "troe":
            lines += ["  type: falloff", f"  low-P-rate-constant: {_map(x['low-P-rate-constant'])}",
                      f"  high-P-rate-constant: {_map(x['rate-constant'])}", f"  Troe: {_map(x['Troe'])}"]
        else:
            lines.append(f"  rate-constant: {_map(x['rate-constant'])}")
        if "efficiencies" in x:
            lines.append(f"  efficiencies: {_map(x['efficiencies'])}")
    return "\n".join(lines) + "\n"


# Named configurations (species, reactions, three-body, Troe, seed).
CONFIGS = {
    "h2o2-10": (10, 21, 4, 1, 10),
    "gri-53": (53, 325, 30, 25, 53),
    "large-160": (160, 1800, 120, 90, 160),
}
# State draws per configuration: T range [K], P range [atm], Dirichlet concentration.
STATE_RANGES = {
    "h2o2-10": ((800.0, 2500.0), (1.0, 10.0), 0.5),
    "gri-53": ((600.0, 2800.0), (1.0, 40.0), 0.3),
    "large-160": ((600.0, 2800.0), (1.0, 40.0), 0.3),
}


def config(name):
    K, R, tb, troe, seed = CONFIGS[name]
    return generate(K, R, tb, troe, seed, name)


def mechanism_path(name, directory=None):
    """Writes (once) and returns the YAML path of a named configuration."""
    directory = directory or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
    path = os.path.join(directory, name + ".yaml")
    if not os.path.exists(path):
        os.makedirs(directory, exist_ok=True)
        text = config(name).yaml()
        tmp = path + f".tmp{os.getpid()}"
        with open(tmp, "w") as f:
            f.write(text)
        os.replace(tmp, path)
    return path


def states(K, n, seed, T_range=(600.0, 2800.0), P_range_atm=(1.0, 40.0), alpha=0.3):
    """SoA state arrays: T[n] (K), P[n] (Pa), Y[K][n] (Dirichlet(alpha) mass fractions)."""
    rng = _rng(seed)
    T = rng.uniform(*T_range, size=n)
    P = rng.uniform(*P_range_atm, size=n) * 101325.0
    Y = np.ascontiguousarray(rng.dirichlet(np.full(K, alpha), size=n).T)
    return T, P, Y


def config_states(name, n, seed):
    (T_range, P_range, alpha) = STATE_RANGES[name]
    return states(CONFIGS[name][0], n, seed, T_range, P_range, alpha)


def sri_variant(name, seed=5, params=None):
    """YAML text of a named synthetic mechanism with every Troe block replaced by an SRI block: drawn
    (a ~ U[0.5, 2], b ~ U[100, 1000] K, c ~ U[500, 3000] K, d ~ U[0.8, 1.2], e ~ U[-0.5, 0.5]; the first one with
    c = 0), or the given (a, b, c, d, e) for all."""
    text = open(mechanism_path(name)).read()
    rng = np.random.default_rng(seed)
    count = [0]

    def block(_):
        if params is not None:
            a, b, c, d, e = params
        else:
            a, b, c, d, e = rng.uniform(0.5, 2), rng.uniform(100, 1000), rng.uniform(500, 3000), rng.uniform(0.8, 1.2), rng.uniform(-0.5, 0.5)
            if count[0] == 0:
                c = 0.0
        count[0] += 1
        return f"SRI: {{A: {a!r}, B: {b!r}, C: {c!r}, D: {d!r}, E: {e!r}}}"
    return re.sub(r"Troe: \{[^}]*\}", block, text)


def plog_variant(name, every=3, seed=6):
    """YAML text of a named synthetic mechanism with every `every`-th elementary reaction turned into a PLOG table:
    entries at 2, 8 and 30 atm (its rate constant scaled by drawn factors and exponents), plus a second entry at
    30 atm that the table sums."""
    text = open(mechanism_path(name)).read()
    rng = np.random.default_rng(seed)
    count = [0]

    def table(m):
        count[0] += 1
        if count[0] % every:
            return m.group(0)
        A, b, Ea = float(m.group(2)), float(m.group(3)), float(m.group(4))
        rows = []
        for P in ("2.0 atm", "8.0 atm", "30.0 atm", "30.0 atm"):
            f, db = rng.uniform(0.2, 5.0), rng.uniform(-0.3, 0.3)
            rows.append(f"  - {{P: {P}, A: {A * f!r}, b: {b + db!r}, Ea: {Ea * rng.uniform(0.8, 1.2)!r}}}")
        return f"{m.group(1)}\n  type: pressure-dependent-Arrhenius\n  rate-constants:\n" + "\n".join(rows)
    return re.sub(r"(- equation: [^\n]*<=>[^\n(]*)\n  rate-constant: \{A: ([^,]*), b: ([^,]*), Ea: ([^}]*)\}", table, text)

def models_variant(name, seed=7):
    """YAML text of a named synthetic mechanism that exercises the rate models the synthetic draws leave out:
    of the falloff reactions, every third loses its Troe block (Lindemann falloff, reaction.rs:153-158) and every
    third from the second gets a constant high-pressure limit (b = Ea = 0: the Troe branch of reaction.rs:172-174);
    every fifth elementary reaction becomes irreversible (`=>`, reaction.rs:272). The seed is kept for the
    signature's sake."""
    del seed
    text = open(mechanism_path(name)).read()
    count = {"falloff": 0, "elementary": 0}

    def falloff(m):
        i = count["falloff"]
        count["falloff"] += 1
        block = m.group(0)
        if i % 3 == 0:
            return re.sub(r"\n  Troe: \{[^}]*\}", "", block)
        if i % 3 == 1:
            block = re.sub(r"high-P-rate-constant: \{A: ([^,]*), b: [^,]*, Ea: [^}]*\}",
                           lambda h: f"high-P-rate-constant: {{A: {h.group(1)}, b: 0.0, Ea: 0.0}}", block)
        return block

    text = re.sub(r"- equation: [^\n]*\(\+M\)[^\n]*\n  type: falloff\n(?:  [^\n]*\n)*?  Troe: \{[^}]*\}", falloff, text)

    def elementary(m):
        count["elementary"] += 1
        return m.group(0).replace(" <=> ", " => ") if count["elementary"] % 5 == 0 else m.group(0)

    return re.sub(r"- equation: [^\n(]*<=>[^\n(]*\n  rate-constant:", elementary, text)


VARIANTS = {"sri": sri_variant, "plog": plog_variant, "models": models_variant}


def variant_text(variant, name):
    """The YAML text of a named variant ("sri", "plog", "models") of a named synthetic mechanism."""
    return VARIANTS[variant](name)
