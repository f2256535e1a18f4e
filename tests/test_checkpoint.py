"""Checkpoint output (dmdgpu.checkpoint): the reference's swarm HDF5 fields
(src/IO/hdf5-io.cpp:40-276) as raw arrays + XDMF, read back bit-exactly."""
import os
import xml.etree.ElementTree as ET

import numpy as np

import dmdgpu.checkpoint as ck
import dmdgpu.lattice as lat


def test_round_trip_is_bit_exact(tmp_path):
    sd = lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=9)
    sd["xdot"] = np.random.default_rng(1).standard_normal(sd["n"])
    ck.write_checkpoint(str(tmp_path), sd, step=3, time=1.5e-9, attributes={"Fxx": 0.0})
    sd2 = dict(sd)
    sd2["q"] = sd["q"] + 0.01
    ck.write_checkpoint(str(tmp_path), sd2, step=7, time=2.5e-9)
    assert ck.list_steps(str(tmp_path)) == [3, 7]
    back, meta = ck.read_checkpoint(str(tmp_path), step=3)
    assert meta["time"] == 1.5e-9 and meta["attributes"]["Fxx"] == 0.0
    for key in ("q", "sigma", "x", "beta", "gamma", "species", "xdot"):
        assert np.array_equal(back[key], sd[key]), key
    assert back["species"].dtype == np.int32
    assert np.array_equal(back["lo"], sd["lo"]) and np.array_equal(back["hi"], sd["hi"])
    assert np.array_equal(back["periodic"], sd["periodic"])
    last, meta = ck.read_checkpoint(str(tmp_path))
    assert meta["step"] == 7 and np.array_equal(last["q"], sd2["q"])


def test_xdmf_describes_the_arrays(tmp_path):
    sd = lat.cluster(2, 1, 1, x_h=(0.0, 0.3), seed=3)
    ck.write_checkpoint(str(tmp_path), sd, step=0)
    root = ET.parse(os.path.join(str(tmp_path), "swarm_000000.xmf")).getroot()
    grid = root.find("Domain/Grid")
    n = sd["n"]
    assert grid.find("Topology").get("NumberOfElements") == str(n)
    geo = grid.find("Geometry/DataItem")
    assert geo.get("Dimensions") == f"{n} 3" and geo.get("Format") == "Binary"
    assert os.path.getsize(os.path.join(str(tmp_path), geo.text)) == n * 3 * 8
    names = {a.get("Name"): a.find("DataItem") for a in grid.findall("Attribute")}
    assert set(names) == {"stdv-q", "molar-fraction", "molar-fraction-dt", "beta", "gamma",
                          "specie", "idx"}
    item = names["specie"]
    assert item.get("NumberType") == "Int" and item.get("Precision") == "4"
    assert os.path.getsize(os.path.join(str(tmp_path), item.text)) == n * 4
