"""Mechanism loader: Cantera-format YAML -> the flat parameter arrays the
kernels take.

The arrays and their units follow the reference's two parameter bundles:

* ``thermoProperty`` (run.jl:71-82), filled by ``initThermo`` (thermo.jl:405-436):
  Ru = gas_constant*1e-3 [J/mol/K]; mw [kg/mol]; coeffs_sep = NASA-7 mid
  temperature; coeffs_lo/hi [7, Ns] (column per species); min/max_temp of the
  phase; viscosity / conductivity polynomials [5, Ns] and binary-diffusion
  polynomials [5, Ns, Ns] in ln T.
* ``reactionProperty`` (run.jl:84-94), filled by ``initReact``
  (reactions.jl:506-576): reaction_type 1 Arrhenius / 2 three-body / 3 Troe
  falloff; sgm = sum(nu_f - nu_r); vf, vr [Ns, Nr]; Arr [3, Nr] = (A, b, Ea/R)
  with A in mol-cm-s units (A_SI * 1e3^(order-1), order counting M for
  three-body reactions); loP [3, Nr] with A_SI * 1e3^order; Troe [4, Nr] =
  (A, T3, T1, T2) with T2 = 1e30 when absent; efficiencies [Ns, Nr] with the
  default efficiency for species that are not listed.

Cantera itself is not used.  Each rate is read from the YAML text and converted
with the same unit algebra Cantera applies, so ``A`` ends up in mol-cm-s units
exactly as the reference's round trip through Cantera's SI values does (up to
one rounding of the SI round trip).

Transport polynomials come from Cantera's fit of its own collision-integral
tables in the reference (thermo.jl:425-428).  That fit is not reproduced
here: ``fit_transport`` builds polynomials of the same form from
Chapman-Enskog theory with the Neufeld-Janzen-Aziz (1972) collision-integral
correlations and a modified-Eucken conductivity.  The kernels consume the
polynomials as data, so GPU-vs-oracle parity does not depend on them, but
their values are not Cantera's (the coefficients are not pinned).
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
import re
from typing import Dict, List, Optional

import numpy as np

# Cantera 3.x constants (ct.gas_constant [J/kmol/K], ct.one_atm [Pa]).
GAS_CONSTANT = 8314.46261815324
ONE_ATM = 101325.0
AVOGADRO = 6.02214076e26  # per kmol

# Atomic weights used by Cantera 3.x (IUPAC abridged values).
ATOMIC_WEIGHTS = {"H": 1.008, "O": 15.999, "N": 14.007, "C": 12.011,
                  "Ar": 39.95, "AR": 39.95, "He": 4.002602, "HE": 4.002602}

_ENERGY_TO_J_PER_KMOL = {
    "cal/mol": 4184.0, "kcal/mol": 4184.0e3, "J/mol": 1.0e3, "kJ/mol": 1.0e6,
    "J/kmol": 1.0, "kJ/kmol": 1.0e3, "cal/kmol": 4.184, "kcal/kmol": 4184.0,
}


@dataclasses.dataclass
class Mechanism:
    """Flat arrays in the reference's layout (Fortran order = column-major)."""
    species: List[str]
    # thermo (thermoProperty)
    Ru: float
    min_temp: float
    max_temp: float
    mw: np.ndarray            # [Ns]
    coeffs_sep: np.ndarray    # [Ns]
    coeffs_lo: np.ndarray     # [7, Ns]
    coeffs_hi: np.ndarray     # [7, Ns]
    visc_poly: np.ndarray     # [5, Ns]
    conduct_poly: np.ndarray  # [5, Ns]
    binarydiff_poly: np.ndarray  # [5, Ns, Ns]
    # kinetics (reactionProperty)
    atm: float
    reaction_type: np.ndarray  # [Nr] int64
    sgm: np.ndarray            # [Nr] int64
    vf: np.ndarray             # [Ns, Nr] int64
    vr: np.ndarray             # [Ns, Nr] int64
    Arr: np.ndarray            # [3, Nr]
    ef: np.ndarray             # [Ns, Nr]
    loP: np.ndarray            # [3, Nr]
    Troe: np.ndarray           # [4, Nr]
    equations: List[str] = dataclasses.field(default_factory=list)
    elements: List[str] = dataclasses.field(default_factory=list)
    composition: Optional[np.ndarray] = None  # [n_elements, Ns] atoms per molecule

    @property
    def Ns(self) -> int:
        return len(self.species)

    @property
    def Nr(self) -> int:
        return int(self.reaction_type.shape[0])

    def index(self, name: str) -> int:
        return self.species.index(name)

    # ---- serialisation (JSON with plain lists; column-major flattening) ----
    def to_json(self) -> str:
        d = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if v is None:
                continue
            if isinstance(v, np.ndarray):
                d[f.name] = {"shape": list(v.shape), "dtype": str(v.dtype),
                             "data": v.flatten(order="F").tolist()}
            else:
                d[f.name] = v
        return json.dumps(d, indent=1)

    @staticmethod
    def from_json(text: str) -> "Mechanism":
        d = json.loads(text)
        kw = {}
        for f in dataclasses.fields(Mechanism):
            if f.name not in d:
                continue
            v = d[f.name]
            if isinstance(v, dict) and "shape" in v:
                kw[f.name] = np.array(v["data"], dtype=v["dtype"]).reshape(
                    v["shape"], order="F")
            else:
                kw[f.name] = v
        return Mechanism(**kw)


def _parse_quantity(val, default_unit: str) -> float:
    """Return an activation energy in J/kmol from a YAML value such as
    ``16.6 kcal/mol`` or a bare number in the file's default unit."""
    if isinstance(val, (int, float)):
        return float(val) * _ENERGY_TO_J_PER_KMOL[default_unit]
    m = re.match(r"^\s*([-+0-9.eE]+)\s*(\S+)\s*$", str(val))
    if not m:
        raise ValueError(f"cannot parse quantity {val!r}")
    unit = m.group(2)
    if unit == "K":
        return float(m.group(1)) * GAS_CONSTANT
    return float(m.group(1)) * _ENERGY_TO_J_PER_KMOL[unit]


def _parse_composition(side: str) -> Dict[str, int]:
    comp: Dict[str, int] = {}
    for term in side.split(" + "):
        term = term.strip()
        if not term or term == "M" or term.startswith("(+"):
            continue
        m = re.match(r"^(\d+)\s+(\S+)$", term)
        if m:
            n, sp = int(m.group(1)), m.group(2)
        else:
            n, sp = 1, term
        comp[sp] = comp.get(sp, 0) + n
    return comp


def _split_equation(eq: str):
    eq = re.sub(r"\(\s*\+\s*M\s*\)", "", eq)
    eq = re.sub(r"\(\+M\)", "", eq)
    for arrow in ("<=>", "=>", "="):
        if arrow in eq:
            lhs, rhs = eq.split(arrow)
            return _parse_composition(lhs), _parse_composition(rhs)
    raise ValueError(eq)


def load_yaml(path: str, fit: bool = True) -> Mechanism:
    """Parse a Cantera YAML mechanism (the subset the reference supports:
    elementary Arrhenius, three-body, Troe falloff; NASA-7 thermo)."""
    import yaml

    class _Loader(yaml.SafeLoader):
        """SafeLoader without YAML 1.1 booleans, so species named NO / N / Y
        stay strings."""
    _Loader.yaml_implicit_resolvers = {
        k: [r for r in v if r[0] != "tag:yaml.org,2002:bool"]
        for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()}
    with open(path) as fh:
        doc = yaml.load(fh, Loader=_Loader)
    units = doc.get("units", {})
    length = units.get("length", "m")
    quantity = units.get("quantity", "kmol")
    ea_unit = units.get("activation-energy", "J/kmol")
    # factor that turns a rate constant of overall order n (in the file's
    # units) into SI (m^3/kmol)^(n-1)/s: (length^3/quantity)^(n-1)
    # (length^3/quantity) expressed in cm^3/mol, the unit the reference's
    # A * 1e3^(order-1) lands in; exact powers of ten, so mol-cm files pass
    # through unchanged.
    cm3_per_mol = {("cm", "mol"): 1.0, ("m", "kmol"): 1e3, ("m", "mol"): 1e6,
                   ("cm", "kmol"): 1e-3, ("mm", "mol"): 1e-3}[(length, quantity)]

    phase = doc["phases"][0]
    names = list(phase["species"])
    sp_by_name = {s["name"]: s for s in doc["species"]}
    Ns = len(names)

    mw = np.zeros(Ns)
    sep = np.zeros(Ns)
    lo = np.zeros((7, Ns))
    hi = np.zeros((7, Ns))
    tmins, tmaxs = [], []
    for n, name in enumerate(names):
        s = sp_by_name[name]
        w = 0.0
        for el, cnt in s["composition"].items():
            w += ATOMIC_WEIGHTS[el] * cnt
        mw[n] = w * 1e-3
        th = s["thermo"]
        if th["model"] != "NASA7":
            raise ValueError("only NASA7 thermo is supported")
        tr = th["temperature-ranges"]
        data = th["data"]
        tmins.append(tr[0])
        tmaxs.append(tr[-1])
        if len(tr) == 2:
            sep[n] = tr[1]
            lo[:, n] = data[0]
            hi[:, n] = data[0]
        else:
            sep[n] = tr[1]
            lo[:, n] = data[0]
            hi[:, n] = data[1]

    reactions = doc.get("reactions", [])
    if isinstance(reactions, str):
        raise ValueError("reaction sections by name are not supported")
    Nr = len(reactions)
    rtype = np.zeros(Nr, dtype=np.int64)
    sgm = np.zeros(Nr, dtype=np.int64)
    vf = np.zeros((Ns, Nr), dtype=np.int64)
    vr = np.zeros((Ns, Nr), dtype=np.int64)
    Arr = np.zeros((3, Nr))
    ef = np.zeros((Ns, Nr))
    loP = np.zeros((3, Nr))
    Troe = np.zeros((4, Nr))
    eqs = []

    def rate_cgs(rc, order):
        """A in mol-cm-s units for a rate of overall order ``order``; equals
        Cantera's SI value times 1e3^(order-1) as initReact forms it."""
        return float(rc["A"]) * cm3_per_mol ** (order - 1)

    for j, r in enumerate(reactions):
        eq = r["equation"]
        eqs.append(eq)
        reac, prod = _split_equation(eq)
        for sp, c in reac.items():
            vf[names.index(sp), j] = c
        for sp, c in prod.items():
            vr[names.index(sp), j] = c
        rsum = int(vf[:, j].sum())
        kind = r.get("type", "elementary")
        if kind == "elementary":
            rtype[j] = 1
            order = rsum
            rc = r["rate-constant"]
            Arr[0, j] = rate_cgs(rc, order)
            Arr[1, j] = float(rc["b"])
            Arr[2, j] = _parse_quantity(rc["Ea"], ea_unit) / GAS_CONSTANT
        elif kind == "three-body":
            rtype[j] = 2
            order = rsum + 1
            rc = r["rate-constant"]
            Arr[0, j] = rate_cgs(rc, order)
            Arr[1, j] = float(rc["b"])
            Arr[2, j] = _parse_quantity(rc["Ea"], ea_unit) / GAS_CONSTANT
            dflt = float(r.get("default-efficiency", 1.0))
            effs = r.get("efficiencies", {})
            for n, name in enumerate(names):
                ef[n, j] = float(effs.get(name, dflt))
        elif kind == "falloff":
            if "Troe" not in r:
                raise ValueError(f"reaction {j + 1}: only Troe falloff supported")
            rtype[j] = 3
            order = rsum
            hr = r["high-P-rate-constant"]
            lr = r["low-P-rate-constant"]
            Arr[0, j] = rate_cgs(hr, order)
            Arr[1, j] = float(hr["b"])
            Arr[2, j] = _parse_quantity(hr["Ea"], ea_unit) / GAS_CONSTANT
            loP[0, j] = rate_cgs(lr, order + 1)
            loP[1, j] = float(lr["b"])
            loP[2, j] = _parse_quantity(lr["Ea"], ea_unit) / GAS_CONSTANT
            t = r["Troe"]
            Troe[0, j] = float(t["A"])
            Troe[1, j] = float(t["T3"])
            Troe[2, j] = float(t["T1"])
            Troe[3, j] = float(t.get("T2", 1e30))
            dflt = float(r.get("default-efficiency", 1.0))
            effs = r.get("efficiencies", {})
            for n, name in enumerate(names):
                ef[n, j] = float(effs.get(name, dflt))
        else:
            raise ValueError(f"Not supported reaction type of reaction {j + 1}")
        sgm[j] = int((vf[:, j] - vr[:, j]).sum())

    elements = [e for e in phase.get("elements", []) if any(
        e in sp_by_name[n]["composition"] for n in names)]
    comp = np.zeros((len(elements), Ns))
    for n, name in enumerate(names):
        for el, cnt in sp_by_name[name]["composition"].items():
            comp[elements.index(el), n] = cnt
    mech = Mechanism(
        species=names, Ru=GAS_CONSTANT * 1e-3, elements=elements, composition=comp, min_temp=float(max(tmins)),
        max_temp=float(min(tmaxs)), mw=mw, coeffs_sep=sep, coeffs_lo=lo,
        coeffs_hi=hi, visc_poly=np.zeros((5, Ns)), conduct_poly=np.zeros((5, Ns)),
        binarydiff_poly=np.zeros((5, Ns, Ns)), atm=ONE_ATM, reaction_type=rtype,
        sgm=sgm, vf=vf, vr=vr, Arr=Arr, ef=ef, loP=loP, Troe=Troe, equations=eqs)
    if fit:
        trans = {}
        for name in names:
            t = sp_by_name[name].get("transport")
            if t is None:
                raise ValueError(f"species {name} has no transport data")
            trans[name] = t
        fit_transport(mech, trans)
    return mech


# ---------------------------------------------------------------------------
# Transport polynomials (Chapman-Enskog + Neufeld et al. 1972 correlations)
# ---------------------------------------------------------------------------
_KB = 1.380649e-23


def _omega22(ts):
    return 1.16145 * ts ** -0.14874 + 0.52487 * np.exp(-0.77320 * ts) + \
        2.16178 * np.exp(-2.43787 * ts)


def _omega11(ts):
    return 1.06036 * ts ** -0.15610 + 0.19300 * np.exp(-0.47635 * ts) + \
        1.03587 * np.exp(-1.52996 * ts) + 1.76474 * np.exp(-3.89411 * ts)


def _cp_R(mech: Mechanism, n: int, T: np.ndarray) -> np.ndarray:
    c = np.where(T[None, :] < mech.coeffs_sep[n], mech.coeffs_lo[:5, n:n + 1],
                 mech.coeffs_hi[:5, n:n + 1])
    return c[0] + c[1] * T + c[2] * T ** 2 + c[3] * T ** 3 + c[4] * T ** 4


def fit_transport(mech: Mechanism, trans: Dict[str, dict], npts: int = 60):
    """Fill visc/conduct/binarydiff polynomials in the reference's forms:
    mu = (T^1/4 P(lnT))^2, lambda = sqrt(T) P(lnT), D_kl*P = T^1.5 P(lnT)
    (thermo.jl:289-369)."""
    T = np.linspace(mech.min_temp, mech.max_temp, npts)
    lnT = np.log(T)
    V = np.vander(lnT, 5, increasing=True)
    Ns = mech.Ns
    eps_k = np.array([float(trans[s]["well-depth"]) for s in mech.species])
    sigma = np.array([float(trans[s]["diameter"]) for s in mech.species])
    Wg = mech.mw * 1e3  # g/mol
    Rg = mech.Ru / mech.mw
    for n in range(Ns):
        ts = T / eps_k[n]
        mu = 2.6693e-6 * np.sqrt(Wg[n] * T) / (sigma[n] ** 2 * _omega22(ts))
        cp = _cp_R(mech, n, T) * Rg[n]
        lam = mu * (cp + 1.25 * Rg[n])
        mech.visc_poly[:, n] = np.linalg.lstsq(V, np.sqrt(mu / np.sqrt(T)), rcond=None)[0]
        mech.conduct_poly[:, n] = np.linalg.lstsq(V, lam / np.sqrt(T), rcond=None)[0]
    for n in range(Ns):
        for l in range(Ns):
            ekl = math.sqrt(eps_k[n] * eps_k[l])
            skl = 0.5 * (sigma[n] + sigma[l])
            wkl = 1.0 / Wg[n] + 1.0 / Wg[l]
            ts = T / ekl
            # D [cm^2/s] at P [atm]; D*P in SI = D*1e-4 * P*101325
            dp = 0.0018583 * np.sqrt(T ** 3 * wkl) / (skl ** 2 * _omega11(ts))
            dp_si = dp * 1e-4 * ONE_ATM
            mech.binarydiff_poly[:, n, l] = np.linalg.lstsq(V, dp_si / T ** 1.5, rcond=None)[0]
    return mech


def perfect_gas(R_gas: float = 287.0, gamma: float = 1.4, mu: float = 1.8e-5,
                Pr: float = 0.72, T_lo: float = 200.0, T_hi: float = 400.0) -> Mechanism:
    """One-species calorically perfect gas for the Taylor-Green case: NASA-7
    a1 = gamma/(gamma-1), others zero; constant mu and lambda = cp*mu/Pr
    encoded in the reference's polynomial forms by a least-squares fit over
    [T_lo, T_hi]."""
    a1 = gamma / (gamma - 1.0)
    W = 8.31446261815324 / R_gas
    lo = np.zeros((7, 1))
    lo[0, 0] = a1
    T = np.linspace(T_lo, T_hi, 64)
    V = np.vander(np.log(T), 5, increasing=True)
    cp = a1 * R_gas
    visc = np.linalg.lstsq(V, np.sqrt(mu / np.sqrt(T)), rcond=None)[0]
    cond = np.linalg.lstsq(V, (cp * mu / Pr) / np.sqrt(T), rcond=None)[0]
    return Mechanism(
        species=["GAS"], Ru=GAS_CONSTANT * 1e-3, min_temp=T_lo, max_temp=T_hi,
        mw=np.array([W]), coeffs_sep=np.array([1000.0]), coeffs_lo=lo,
        coeffs_hi=lo.copy(), visc_poly=visc.reshape(5, 1),
        conduct_poly=cond.reshape(5, 1), binarydiff_poly=np.ones((5, 1, 1)) * 1e-5,
        atm=ONE_ATM, reaction_type=np.zeros(0, dtype=np.int64),
        sgm=np.zeros(0, dtype=np.int64), vf=np.zeros((1, 0), dtype=np.int64),
        vr=np.zeros((1, 0), dtype=np.int64), Arr=np.zeros((3, 0)),
        ef=np.zeros((1, 0)), loP=np.zeros((3, 0)), Troe=np.zeros((4, 0)))


_HERE = os.path.dirname(os.path.abspath(__file__))
MECH_DIR = os.path.join(_HERE, "mechanisms")


def load(name: str) -> Mechanism:
    """Load a packaged mechanism: 'air5' (Park 5-species air) or
    'h2_lidryer' (Li-Dryer H2/O2, 9 species)."""
    with open(os.path.join(MECH_DIR, name + ".json")) as fh:
        return Mechanism.from_json(fh.read())
