"""Material registry: named electric/magnetic material pairs and their XML form.

Mirrors the reference's MaterialMap / MaterialPair (include/MaterialMap.hpp:45-284):
a map from an identifier to a pair of material models, one for the electric
and one for the magnetic field, written by ``print_summary`` and read back by
``readXML`` (MaterialPair::readXML 135-170, MaterialMap::readXML 230-266; the
per-model tags of DispersiveMaterials.hpp: Constant 411-430, Conductive
639-660, Lorentz 822-860, Drude 1117-1140, Vacuum 466-470).  The reference
reads XML through the optional tinyxml2; here the standard library's parser
reads the same schema, and ``to_xml`` writes the text ``print_summary``
writes (tabs, order by identifier, C++ default float format).

A map is put on a grid (``fdtd_amd.Grid``) or the oracle by ``apply(sim, ids)``
with ids = {identifier: material id}: electric Vacuum / Constant / Conductive
/ Drude / Lorentz become the material's eps_r and pole, MagnetizedDrude
(MagnetizedDrudeStorage 1619-1726, its <B>(x, y, z)</B> tag included) and
SelfMagnetizedDrude (1898-2006) a magnetized Drude material, FluidDrude
(1160-1199), MagnetizedFluidDrude (1728-1814) and SelfMagnetizedFluidDrude
(1816-1893) a fluid Drude material (its density and collision rate per cell
come from the grid's fluid state), magnetic Vacuum / Constant its mu_r,
magnetic Drude / Lorentz / Conductive its mu_r and magnetic pole (B, H, M, K:
``set_material_magnetic``).  Magnetized and fluid models on the magnetic side
raise.
"""
import xml.etree.ElementTree as ET
from dataclasses import dataclass

from .grid import CONDUCTIVE, DRUDE, LORENTZ

# Detail::Dispersion names (DispersiveMaterials.hpp:2038-2068)
DISPERSIONS = ("Vacuum", "Constant", "Conductive", "Drude", "FluidDrude", "Lorentz", "MagnetizedDrude",
               "SelfMagnetizedDrude", "MagnetizedFluidDrude", "SelfMagnetizedFluidDrude", "MultiCoefficient",
               "Anisotropic")
# the tags each model writes, in print_summary order
_FIELDS = {"Vacuum": (), "Constant": ("Constant",), "Conductive": ("Constant", "Freq"),
           "Drude": ("Constant", "Freq", "Gamma"), "Lorentz": ("Constant", "Delta", "Freq", "Gamma"),
           "MagnetizedDrude": ("B", "Constant", "Freq", "Gamma"),
           "SelfMagnetizedDrude": ("Constant", "Freq", "Gamma"),
           "FluidDrude": (), "MagnetizedFluidDrude": ("B",)}
# models written as one empty element on one line (their print_summary)
_ONE_LINE = ("Vacuum", "SelfMagnetizedFluidDrude")
_FLUID = {"FluidDrude": 0, "MagnetizedFluidDrude": 1, "SelfMagnetizedFluidDrude": 2}
_ATTR = {"Constant": "K", "Freq": "freq", "Gamma": "gamma", "Delta": "delta", "B": "B"}


@dataclass
class Dispersion:
    """One material model: kind (a DISPERSIONS name) and its coefficients (K the
    relative constant, freq, gamma, delta as the reference names them; B the
    bias field [T] of a MagnetizedDrude model)."""
    kind: str
    K: float = 1.0
    freq: float = 0.0
    gamma: float = 0.0
    delta: float = 0.0
    B: tuple = (0.0, 0.0, 0.0)

    def __post_init__(self):
        if self.kind not in DISPERSIONS:
            raise ValueError(f"unknown material model {self.kind!r}")


@dataclass
class MaterialPair:
    electric: Dispersion
    magnetic: Dispersion

    def spec(self):
        """(eps_r, mu_r, poles) for Grid.set_material / the oracle's set_material;
        for a magnetized Drude model (eps_r, mu_r, poles, bias) with bias None for
        the self-magnetized one (Grid.set_magnetized_drude); for a fluid Drude
        model {"fluid": magnetization, "mu": mu_r, "bias": B or None}
        (Grid.set_fluid_drude)."""
        e, m = self.electric, self.magnetic
        bias = ()
        fluid = None
        if e.kind == "Vacuum":
            eps, poles = 1.0, []
        elif e.kind == "Constant":
            eps, poles = e.K, []
        elif e.kind == "Conductive":
            eps, poles = e.K, [(CONDUCTIVE, e.freq, 0.0, 1.0)]
        elif e.kind == "Drude":
            eps, poles = e.K, [(DRUDE, e.freq, e.gamma, 1.0)]
        elif e.kind == "Lorentz":  # the reference's update never reads Delta
            eps, poles = e.K, [(LORENTZ, e.freq, e.gamma, 1.0)]
        elif e.kind in ("MagnetizedDrude", "SelfMagnetizedDrude"):
            eps, poles = e.K, [(DRUDE, e.freq, e.gamma, 1.0)]
            bias = (tuple(float(b) for b in e.B) if e.kind == "MagnetizedDrude" else None,)
        elif e.kind in _FLUID:
            eps, poles = 1.0, [(DRUDE, 0.0, 0.0, 1.0)]
            fluid = _FLUID[e.kind]
        else:
            raise NotImplementedError(f"electric model {e.kind} has no device counterpart")
        if m.kind == "Vacuum":
            mu = 1.0
        elif m.kind in ("Constant", "Drude", "Lorentz", "Conductive"):
            mu = m.K  # poles of the last three: magnetic_poles()
        else:
            raise NotImplementedError(f"magnetic model {m.kind} has no device counterpart")
        if fluid is not None:
            return {"fluid": fluid, "mu": mu, "bias": tuple(float(b) for b in e.B) if fluid == 1 else None}
        return (eps, mu, poles) + bias

    def magnetic_poles(self):
        """Poles of the magnetic model (the reference's Drude / Lorentz / Conductive
        updates on FieldType::Magnetic, stepping B, H, M, K) for set_material_magnetic;
        [] for Vacuum and Constant."""
        m = self.magnetic
        if m.kind == "Conductive":
            return [(CONDUCTIVE, m.freq, 0.0, 1.0)]
        if m.kind == "Drude":
            return [(DRUDE, m.freq, m.gamma, 1.0)]
        if m.kind == "Lorentz":
            return [(LORENTZ, m.freq, m.gamma, 1.0)]
        return []


def _fmt(v):
    """std::ostream << double with the default flags (precision 6, %g)."""
    return format(float(v), "g")


class MaterialMap(dict):
    """identifier -> MaterialPair (std::map: written in identifier order)."""

    @classmethod
    def read_xml(cls, source):
        """Parse a <MaterialMap> document (a path, or the text itself if it starts with '<')."""
        text = source if source.lstrip().startswith("<") else open(source).read()
        root = ET.fromstring(text)
        if root.tag != "MaterialMap":
            raise ValueError(f"expected <MaterialMap>, got <{root.tag}>")
        out = cls()
        for item in root.findall("Item"):
            ident = item.findtext("Identifier")
            pair = item.find("MaterialPair")
            if ident is None or pair is None:
                raise ValueError("an <Item> needs <Identifier> and <MaterialPair>")
            models = {}
            for side in ("Electric", "Magnetic"):
                node = pair.find(side)
                if node is None or len(node) == 0:
                    raise ValueError(f"{ident}: <{side}> needs one model element")
                model = node[0]
                d = Dispersion(model.tag)
                for child in model:
                    if child.tag == "B":  # "(bx, by, bz)", MagnetizedDrudeStorage::readXML
                        d.B = tuple(float(v) for v in child.text.strip().strip("()").split(","))
                    elif child.tag in _ATTR:
                        setattr(d, _ATTR[child.tag], float(child.text))
                models[side] = d
            out[ident] = MaterialPair(models["Electric"], models["Magnetic"])
        return out

    def to_xml(self, ntabs=0):
        """The text MaterialMap::print_summary writes (MaterialMap.hpp:204-228)."""
        t = lambda n: "\t" * n
        lines = [t(ntabs) + "<MaterialMap>"]
        for ident in sorted(self):
            pair = self[ident]
            lines += [t(ntabs + 1) + "<Item>", t(ntabs + 2) + f"<Identifier>{ident}</Identifier>",
                      t(ntabs + 2) + "<MaterialPair>"]
            for side, d in (("Electric", pair.electric), ("Magnetic", pair.magnetic)):
                lines.append(t(ntabs + 3) + f"<{side}>")
                if d.kind in _ONE_LINE:
                    lines.append(t(ntabs + 4) + f"<{d.kind}></{d.kind}>")
                else:
                    if d.kind not in _FIELDS:
                        raise NotImplementedError(f"writing model {d.kind}")
                    lines.append(t(ntabs + 4) + f"<{d.kind}>")
                    for tag in _FIELDS[d.kind]:
                        if tag == "B":
                            v = "(" + ", ".join(_fmt(b) for b in d.B) + ")"
                        else:
                            v = _fmt(getattr(d, _ATTR[tag]))
                        lines.append(t(ntabs + 5) + f"<{tag}>{v}</{tag}>")
                    lines.append(t(ntabs + 4) + f"</{d.kind}>")
                lines.append(t(ntabs + 3) + f"</{side}>")
            lines += [t(ntabs + 2) + "</MaterialPair>", t(ntabs + 1) + "</Item>"]
        lines.append(t(ntabs) + "</MaterialMap>")
        return "\n".join(lines) + "\n"

    def apply(self, sim, ids):
        """Set material id ids[name] of `sim` (Grid or oracle Sim) from each named pair."""
        for name, mid in ids.items():
            s = self[name].spec()
            if isinstance(s, dict):
                sim.set_fluid_drude(mid, s["mu"], s["fluid"], s["bias"])
            elif len(s) == 4:
                eps, mu, poles, bias = s
                sim.set_magnetized_drude(mid, eps, mu, poles[0][1], poles[0][2], bias)
            else:
                sim.set_material(mid, *s)
            mp = self[name].magnetic_poles()
            if mp:
                sim.set_material_magnetic(mid, mp)
        return sim
