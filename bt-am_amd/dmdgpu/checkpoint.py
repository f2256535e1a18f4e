"""Swarm checkpoints: output for visual checks and resume.

The reference writes one HDF5 file per output step from its DMSwarm
(src/IO/hdf5-io.cpp:40-276): the box (Xmin..Zmax) and load attributes, and
the site fields coordinates, stdv-q, molar-fraction, molar-fraction-dt, beta,
gamma, idx and specie; it reads them back to restart (:280-504).  HDF5 is not
available here, so a checkpoint is a directory holding each field as a raw
little-endian array, a JSON file with the attributes and the array layout, and
an XDMF 3 description (Polyvertex topology, XYZ geometry, one attribute per
field, Format="Binary") that ParaView and VisIt read directly.
"""
import json
import os
import xml.etree.ElementTree as ET

import numpy as np

# (field name in the reference swarm, key in a system dict, dtype, components)
FIELDS = [
    ("coordinates", "q", "<f8", 3),
    ("stdv-q", "sigma", "<f8", 1),
    ("molar-fraction", "x", "<f8", 1),
    ("molar-fraction-dt", "xdot", "<f8", 1),
    ("beta", "beta", "<f8", 1),
    ("gamma", "gamma", "<f8", 1),
    ("specie", "species", "<i4", 1),
    ("idx", "idx", "<i4", 1),
]


def _xdmf(meta, n):
    root = ET.Element("Xdmf", Version="3.0")
    grid = ET.SubElement(ET.SubElement(root, "Domain"), "Grid", Name="swarm",
                         GridType="Uniform")
    ET.SubElement(grid, "Time", Value=repr(meta["time"]))
    ET.SubElement(grid, "Topology", TopologyType="Polyvertex", NumberOfElements=str(n))
    for name, f in meta["fields"].items():
        prec = "8" if f["dtype"] == "<f8" else "4"
        num = "Float" if f["dtype"] == "<f8" else "Int"
        dims = f"{n} {f['components']}" if f["components"] > 1 else str(n)
        if name == "coordinates":
            parent = ET.SubElement(grid, "Geometry", GeometryType="XYZ")
        else:
            parent = ET.SubElement(grid, "Attribute", Name=name, Center="Node",
                                   AttributeType="Vector" if f["components"] == 3 else "Scalar")
        item = ET.SubElement(parent, "DataItem", Dimensions=dims, NumberType=num,
                             Precision=prec, Format="Binary", Endian="Little")
        item.text = f["file"]
    return ET.ElementTree(root)


def write_checkpoint(folder, sysd, step=0, time=0.0, attributes=None):
    """Write the system dict ``sysd`` (lattice layout: n, q, sigma, x, species,
    beta, gamma, lo, hi, periodic; optional xdot) as output step ``step``.
    Returns the path of the step's directory; ``swarm_<step>.xmf`` beside it
    describes the same arrays."""
    n = int(sysd["n"])
    name = f"step_{int(step):06d}"
    d = os.path.join(folder, name)
    os.makedirs(d, exist_ok=True)
    src = dict(sysd)
    src.setdefault("xdot", np.zeros(n))
    src.setdefault("idx", np.arange(n, dtype=np.int32))
    meta = {"n": n, "step": int(step), "time": float(time),
            "box": {"lo": [float(v) for v in sysd["lo"]], "hi": [float(v) for v in sysd["hi"]],
                    "periodic": [int(v) for v in sysd["periodic"]]},
            "attributes": dict(attributes or {}), "fields": {}}
    for fname, key, dt, comp in FIELDS:
        a = np.ascontiguousarray(np.asarray(src[key]).astype(dt, copy=False))
        if a.size != n * comp:
            raise ValueError(f"field {fname}: {a.size} values for {n} sites x {comp}")
        rel = f"{name}/{fname}.bin"
        a.tofile(os.path.join(folder, rel))
        meta["fields"][fname] = {"file": rel, "dtype": dt, "components": comp}
    with open(os.path.join(d, "meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    _xdmf(meta, n).write(os.path.join(folder, f"swarm_{int(step):06d}.xmf"),
                         encoding="utf-8", xml_declaration=True)
    return d


def list_steps(folder):
    return sorted(int(e[5:]) for e in os.listdir(folder)
                  if e.startswith("step_") and os.path.isdir(os.path.join(folder, e)))


def read_checkpoint(folder, step=None):
    """Read output step ``step`` (the last one when None) back into a system
    dict; also returns its meta data (time, attributes)."""
    steps = list_steps(folder)
    if not steps:
        raise FileNotFoundError(f"no checkpoint in {folder}")
    step = steps[-1] if step is None else int(step)
    with open(os.path.join(folder, f"step_{step:06d}", "meta.json")) as fh:
        meta = json.load(fh)
    n = meta["n"]
    out = {"n": n, "lo": np.array(meta["box"]["lo"]), "hi": np.array(meta["box"]["hi"]),
           "periodic": np.array(meta["box"]["periodic"], dtype=np.int32)}
    for fname, key, dt, comp in FIELDS:
        f = meta["fields"][fname]
        a = np.fromfile(os.path.join(folder, f["file"]), dtype=f["dtype"])
        if a.size != n * comp:
            raise ValueError(f"field {fname}: {a.size} values, expected {n * comp}")
        out[key] = a.reshape(n, comp) if comp > 1 else a
    out["species"] = out["species"].astype(np.int32)
    return out, meta
