"""Transport fits of a gas mixture -> the tables of include/lsm_transport.h.

The fits are those OpenSMOKE++'s TransportPropertiesMap_CHEMKIN holds
(maps/TransportPropertiesMap_CHEMKIN.hpp:288-403): per species a molecular weight
and cubic fits in ln T of ln lambda and ln eta, per pair j < k a cubic fit in ln T of
ln D_jk at 1 bar (the map keeps 1/D_jk = P_bar/exp(fit)), and for the species with
MW <= 20 (PhysicalConstants.h:44) cubic fits in T of the thermal diffusion ratios
theta_kj.

* :func:`read_transport_xml` reads the ``<Transport type="CHEMKIN">`` block of a
  kinetics.xml (ImportCoefficientsFromXMLFile / ImportCoefficientsFromASCIIFile,
  :340-403, 420-465) and the optional ``<ViscosityModel>`` (:634-651);
  :func:`write_transport_block` writes that block.
* :class:`TransportOptions` holds the solver's physical-model options of the property
  refresh (readOptions.H:94-181, 283-360): Lewis-number diffusivities, simplified
  transport properties, Soret coefficients and the soot-diffusivity corrections that
  :func:`soot_diffusivity_corrections` derives from the BIN species.
* :func:`synthetic_transport` makes physically shaped fits for any mechanism
  (Chapman-Enskog viscosity and diffusion with Neufeld's collision-integral
  correlations, an Eucken-type conductivity) for tests and benchmarks; it is not the
  reference's preprocessor, which fits the kinetic theory from .TRC data.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._abi import LsmTransport, LsmTransportOptions

VISCOSITY_WILKE = 0
VISCOSITY_HERNING = 1
VISCOSITY_MATHUR_SAXENA = 2
DIFFUSIVITY_MULTICOMPONENT = 0
DIFFUSIVITY_LEWIS_NUMBERS = 1
_DIFFUSIVITY = {"multi-component": DIFFUSIVITY_MULTICOMPONENT, "lewis-numbers": DIFFUSIVITY_LEWIS_NUMBERS}
MAX_MW_THERMALDIFFUSION_RATIOS = 20.0        # math/PhysicalConstants.h:44
_MODELS = {"Wilke": VISCOSITY_WILKE, "Herning": VISCOSITY_HERNING, "MathurSaxena": VISCOSITY_MATHUR_SAXENA}


class TransportXmlError(ValueError):
    pass


@dataclass
class Transport:
    """Flat transport fits (0-based, SI).  Field meaning: include/lsm_transport.h."""

    mw: np.ndarray                  # [ns]
    lambda_fit: np.ndarray          # [ns, 4]
    eta_fit: np.ndarray             # [ns, 4]
    gamma_fit: np.ndarray           # [ns(ns-1)/2, 4]
    teta_sp: np.ndarray             # [nteta]
    teta_fit: np.ndarray            # [nteta, ns, 4]
    viscosity_model: int = VISCOSITY_WILKE
    _struct: object = field(default=None, repr=False)

    def __post_init__(self):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.mw = f64(self.mw)
        ns = self.mw.shape[0]
        self.lambda_fit = f64(self.lambda_fit).reshape(ns, 4)
        self.eta_fit = f64(self.eta_fit).reshape(ns, 4)
        self.gamma_fit = f64(self.gamma_fit).reshape(ns * (ns - 1) // 2, 4)
        self.teta_sp = np.ascontiguousarray(self.teta_sp, dtype=np.int32)
        self.teta_fit = f64(self.teta_fit).reshape(len(self.teta_sp), ns, 4)

    @property
    def ns(self) -> int:
        return int(self.mw.shape[0])

    def struct(self) -> LsmTransport:
        if self._struct is None:
            P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
            d, i = ctypes.c_double, ctypes.c_int32
            s = LsmTransport()
            s.ns = self.ns
            s.mw, s.lambda_fit, s.eta_fit = P(self.mw, d), P(self.lambda_fit, d), P(self.eta_fit, d)
            s.gamma_fit = P(self.gamma_fit, d)
            s.nteta = len(self.teta_sp)
            s.teta_sp, s.teta_fit = P(self.teta_sp, i), P(self.teta_fit, d)
            s.viscosity_model = int(self.viscosity_model)
            self._struct = s
        return self._struct


@dataclass
class TransportOptions:
    """Options of the property refresh (include/lsm_transport.h, lsm_transport_options).

    diffusivity: "multi-component" or "lewis-numbers" (readOptions.H:161-181), with
    `lewis` [ns]; simplified: simplifiedTransportProperties; soret: soretEffect (the
    thermal-diffusion output then holds Dsoret); soot_reference / soot_species /
    soot_correction: physicalSootDiffusivity (see soot_diffusivity_corrections)."""

    diffusivity: str = "multi-component"
    lewis: np.ndarray | None = None
    simplified: bool = False
    soret: bool = False
    soot_reference: int = -1
    soot_species: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    soot_correction: np.ndarray = field(default_factory=lambda: np.zeros(0))
    _struct: object = field(default=None, repr=False)

    def __post_init__(self):
        if self.diffusivity not in _DIFFUSIVITY:
            raise ValueError(f"diffusivity must be one of {sorted(_DIFFUSIVITY)}")
        if self.lewis is not None:
            self.lewis = np.ascontiguousarray(self.lewis, dtype=np.float64)
        self.soot_species = np.ascontiguousarray(self.soot_species, dtype=np.int32)
        self.soot_correction = np.ascontiguousarray(self.soot_correction, dtype=np.float64)

    def struct(self) -> LsmTransportOptions:
        if self._struct is None:
            P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t)) if a is not None and a.size else None
            s = LsmTransportOptions()
            s.diffusivity_model = _DIFFUSIVITY[self.diffusivity]
            s.lewis = P(self.lewis, ctypes.c_double)
            s.simplified, s.soret = int(bool(self.simplified)), int(bool(self.soret))
            s.soot_reference = int(self.soot_reference)
            s.n_soot = int(self.soot_species.size)
            s.soot_sp, s.soot_corr = P(self.soot_species, ctypes.c_int32), P(self.soot_correction, ctypes.c_double)
            self._struct = s
        return self._struct


def soot_diffusivity_corrections(species, mw, index: int, exponent: float = -0.681, additional: float = 1.0):
    """(reference, species, corrections) of physicalSootDiffusivity (readOptions.H:318-360):
    the reference is BIN1A; every species named BIN* gets
    min(MW_i / MW_ref, 2^index)^exponent, times `additional` when MW_i / MW_ref >= 10."""
    names = list(species)
    if "BIN1A" not in names:
        raise ValueError("Species BIN1A is not included in the kinetic mechanism")
    ref = names.index("BIN1A")
    sp, corr = [], []
    for i, name in enumerate(names):
        if name[:3] == "BIN":
            r = mw[i] / mw[ref]
            c = min(r, 2.0 ** float(index)) ** exponent
            corr.append(c if r < 10.0 else c * additional)
            sp.append(i)
    return ref, np.array(sp, dtype=np.int32), np.array(corr)


def read_transport_xml(text: str, ns: int | None = None) -> Transport:
    """The transport fits of a kinetics.xml document (the text of the whole file)."""
    import xml.etree.ElementTree as ET
    root = ET.fromstring(text)
    if root.tag != "opensmoke":
        root = root.find("opensmoke")
    if ns is None:
        ns = int(root.find("NumberOfSpecies").text.split()[0])
    node = root.find("Transport")
    if node is None:
        raise TransportXmlError("Transport tag was not found")
    if node.get("type") != "CHEMKIN":
        raise TransportXmlError("Transport type not supported")
    tok = node.text.split()
    pos = [0]

    def nxt():
        if pos[0] >= len(tok):
            raise TransportXmlError("unexpected end of the Transport block")
        pos[0] += 1
        return tok[pos[0] - 1]

    nteta = int(nxt())
    mw, lam, eta, gam, teta_sp, teta = [], [], [], [], [], []
    for j in range(ns):
        mw.append(float(nxt()))
        lam.append([float(nxt()) for _ in range(4)])
        eta.append([float(nxt()) for _ in range(4)])
        for _ in range(j + 1, ns):
            gam.append([float(nxt()) for _ in range(4)])
        if mw[-1] <= MAX_MW_THERMALDIFFUSION_RATIOS:
            teta_sp.append(j)
            teta.append([[float(nxt()) for _ in range(4)] for _ in range(ns)])
    if nxt() != "E" or len(teta_sp) != nteta:
        raise TransportXmlError("Transport block does not end with E after the stated species")
    vm = root.find("ViscosityModel")
    model = VISCOSITY_WILKE
    if vm is not None:
        name = vm.text.strip()
        if name not in _MODELS:
            raise TransportXmlError(f"unknown viscosity model {name!r}")
        model = _MODELS[name]
    return Transport(mw=np.array(mw), lambda_fit=np.array(lam), eta_fit=np.array(eta),
                     gamma_fit=np.array(gam).reshape(-1, 4), teta_sp=np.array(teta_sp, dtype=np.int32),
                     teta_fit=np.array(teta).reshape(len(teta_sp), ns, 4), viscosity_model=model)


def write_transport_block(tr: Transport) -> str:
    """The ``<Transport type="CHEMKIN">`` block (and ``<ViscosityModel>`` when not Wilke)."""
    ns = tr.ns
    teta_of = {int(k): q for q, k in enumerate(tr.teta_sp)}
    out = [str(len(tr.teta_sp))]
    q = 0
    for j in range(ns):
        vals = [tr.mw[j], *tr.lambda_fit[j], *tr.eta_fit[j]]
        for _ in range(j + 1, ns):
            vals += list(tr.gamma_fit[q])
            q += 1
        if j in teta_of:
            vals += list(tr.teta_fit[teta_of[j]].ravel())
        out.append(" ".join(f"{v:.17g}" for v in vals))
    out.append("E")
    text = '<Transport type="CHEMKIN">\n' + "\n".join(out) + "\n</Transport>"
    if tr.viscosity_model != VISCOSITY_WILKE:
        name = {v: k for k, v in _MODELS.items()}[tr.viscosity_model]
        text += f"\n<ViscosityModel> {name} </ViscosityModel>"
    return text


def _omega22(ts):          # Neufeld, Janzen, Aziz, J. Chem. Phys. 57 (1972) 1100
    return 1.16145 * ts ** -0.14874 + 0.52487 * np.exp(-0.77320 * ts) + 2.16178 * np.exp(-2.43787 * ts)


def _omega11(ts):
    return (1.06036 * ts ** -0.15610 + 0.19300 * np.exp(-0.47635 * ts) + 1.03587 * np.exp(-1.52996 * ts) +
            1.76474 * np.exp(-3.89411 * ts))


def synthetic_transport(mech, seed: int = 0, model: int = VISCOSITY_WILKE) -> Transport:
    """Physically shaped fits over 300-3500 K for a mechanism (seeded Lennard-Jones
    parameters); molecular weights from the mechanism."""
    rng = np.random.default_rng(seed)
    ns = mech.ns
    mw = mech.mw
    sigma = 2.5 + 2.5 * rng.random(ns)                  # Angstrom
    eps = 40.0 + 460.0 * rng.random(ns)                 # K
    T = np.linspace(300.0, 3500.0, 64)
    lt = np.log(T)
    fit = lambda v: np.polynomial.polynomial.polyfit(lt, v, 3)
    eta_fit, lam_fit = [], []
    for k in range(ns):
        eta = 2.6693e-6 * np.sqrt(mw[k] * T) / (sigma[k] ** 2 * _omega22(T / eps[k]))
        lam = eta * (8314.4621 / mw[k]) * (1.9 + 1.2 * rng.random())
        eta_fit.append(fit(np.log(eta)))
        lam_fit.append(fit(np.log(lam)))
    gam = []
    for j in range(ns):
        for k in range(j + 1, ns):
            mjk = 2.0 / (1.0 / mw[j] + 1.0 / mw[k])
            sjk = 0.5 * (sigma[j] + sigma[k])
            ejk = np.sqrt(eps[j] * eps[k])
            D = 0.0266e-4 * T ** 1.5 / (np.sqrt(mjk) * sjk ** 2 * _omega11(T / ejk))     # m2/s at 1 bar
            gam.append(fit(np.log(D)))
    teta_sp = [k for k in range(ns) if mw[k] <= MAX_MW_THERMALDIFFUSION_RATIOS]
    teta = []
    for i in teta_sp:
        rows = []
        for j in range(ns):
            th = 0.3 * (mw[j] - mw[i]) / (mw[j] + mw[i]) * (1.0 - 400.0 / T) if j != i else 0.0 * T
            rows.append(np.polynomial.polynomial.polyfit(T, th, 3))
        teta.append(rows)
    return Transport(mw=mw.copy(), lambda_fit=np.array(lam_fit), eta_fit=np.array(eta_fit),
                     gamma_fit=np.array(gam).reshape(-1, 4), teta_sp=np.array(teta_sp, dtype=np.int32),
                     teta_fit=np.array(teta).reshape(len(teta_sp), ns, 4), viscosity_model=model)
