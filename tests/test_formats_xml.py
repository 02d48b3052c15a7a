"""XML restart files in WriteConfig's layout (src/WriteConfig.cpp): the exact
text of a two-atom configuration, written out by hand from the reference's
format strings, in both the text and the base64 encodings, and read-back round
trips of a TIP4P/2005 box with its rigid-water and LJ fix chunks."""
import numpy as np

from md_engine_amd import formats, systems


def two_atoms():
    return dict(turn=12, x=np.array([[0.5, 1.25, -2.0], [3.0, 4.0, 5.0]]), v=np.zeros((2, 3)),
                f=np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]]), type=[0, 1], id=[7, 3], q=[0.5564, -1.1128],
                group_tag=[1, 3], lo=[0.0, 0.0, 0.0], hi=[10.0, 10.0, 10.0], periodic=[1, 1, 0], rcut=9.0,
                padding=1.0, dt=2.0, handles=["O", "H"], masses=[15.9994, 1.008],
                fixes=[("LJCut", "ljcut", formats.lj_restart_chunk([[0.1852, 0], [0, 0]],
                                                                   [[3.1589, 1], [1, 1]], [[9, 9], [9, 9]]))],
                groups={"water": 2, "all": 1}, molecules=[[7, 3]])


def test_text_configuration_exact(tmp_path):
    p = tmp_path / "c.xml"
    formats.write_xml(p, two_atoms())
    want = (
        "<data>\n"
        '<configuration turn="12" numAtoms="2" dimension="3" periodic="110" rCut="9.000000" padding="1.000000" '
        'dt="2.000000">\n'
        '<bounds base64="0" xlo="0.000000" ylo="0.000000" zlo="0.000000" xhi="10.000000" yhi="10.000000" '
        'zhi="10.000000" />\n'
        '<atomParams numTypes="2">\n<handle>\nO\nH\n</handle>\n<mass>\n15.9994\n1.008\n</mass>\n</atomParams>\n'
        "<fixes>\n"
        '<fix type="LJCut" handle="ljcut">\n'
        '<parameter handle="eps">\n0.1852\n0\n0\n0\n</parameter>\n'
        '<parameter handle="rCut">\n9\n9\n9\n9\n</parameter>\n'
        '<parameter handle="sig">\n3.1589\n1\n1\n1\n</parameter>\n'
        "</fix>\n"
        "</fixes>\n"
        "<position>\n0.500000 1.250000 -2.000000\n3.000000 4.000000 5.000000\n</position>\n"
        "<velocity>\n0.000000 0.000000 0.000000\n0.000000 0.000000 0.000000\n</velocity>\n"
        "<force>\n1.000000 0.000000 0.000000\n-1.000000 0.000000 0.000000\n</force>\n"
        "<groupTag>\n1\n3\n</groupTag>\n"
        "<type>\n0\n1\n</type>\n"
        "<id>\n7\n3\n</id>\n"
        "<q>\n0.556400\n-1.112800\n</q>\n"
        "<groupInfo>\n<groupHandles>\nall\nwater\n</groupHandles>\n<groupBits>\n1\n2\n</groupBits>\n</groupInfo>\n"
        "<molecules>\n<m>\n7 3 </m>\n</molecules>\n"
        "</configuration>\n"
        "</data>\n")
    assert p.read_text() == want


def test_base64_configuration_exact(tmp_path):
    p = tmp_path / "c.xml"
    formats.write_xml(p, two_atoms(), base64=True, one_file_per_write=False)
    t = p.read_text()
    # 0.0 and 10.0 as little-endian IEEE doubles; ints as little-endian 32-bit
    assert '<bounds base64="1" xlo="AAAAAAAAAAA=" ylo="AAAAAAAAAAA=" zlo="AAAAAAAAAAA=" xhi="AAAAAAAAJEA=" ' \
           'yhi="AAAAAAAAJEA=" zhi="AAAAAAAAJEA=" />\n' in t
    assert '<type base64="1">\nAAAAAAEAAAA=</type>\n' in t
    assert '<id base64="1">\nBwAAAAMAAAA=</id>\n' in t
    assert "<fixes>" not in t
    assert t.endswith("</configuration>\n</data>")


def water_cfg(box, turn):
    n = len(box["x"])
    nmol = len(box["mol"])
    rng = np.random.default_rng(turn)
    p = systems.TIP4P2005
    return dict(turn=turn, x=box["x"], v=rng.normal(0, 0.01, (n, 3)), f=rng.normal(0, 5, (n, 3)),
                type=box["type"], id=np.arange(n), q=box["q"], group_tag=np.ones(n, np.uint32),
                lo=box["lo"], hi=box["lo"] + box["L"], periodic=[1, 1, 1], rcut=9.0, padding=1.0, dt=2.0,
                handles=["OW", "HW", "M"], masses=[p["m_O"], p["m_H"], p["m_M"]],
                fixes=[("Rigid", "rigid", formats.rigid_restart_chunk(box["mol"]))],
                molecules=[list(m) for m in box["mol"][: nmol // 2]])


def test_round_trip_base64_is_bit_exact(tmp_path):
    box = systems.tip4p2005_box(27, seed=3, min_sep=1.6)
    cfgs = [water_cfg(box, 100), water_cfg(box, 200)]
    p = tmp_path / "w.xml"
    formats.write_xml(p, cfgs, base64=True)
    back = formats.read_xml(p)
    assert [b["turn"] for b in back] == [100, 200]
    for a, b in zip(cfgs, back):
        for k in ("x", "v", "f", "q"):
            assert np.array_equal(np.asarray(a[k], np.float64), b[k]), k
        for k in ("type", "id", "group_tag"):
            assert np.array_equal(np.asarray(a[k]), b[k]), k
        assert np.array_equal(np.asarray(a["hi"], np.float64), b["hi"])
        assert b["molecules"] == [[int(i) for i in m] for m in a["molecules"]]


def test_round_trip_text_to_printed_precision(tmp_path):
    box = systems.tip4p2005_box(27, seed=5, min_sep=1.6)
    cfg = water_cfg(box, 7)
    p = tmp_path / "w.xml"
    formats.write_xml(p, cfg)
    (b,) = formats.read_xml(p)
    for k in ("x", "v", "f", "q"):
        assert np.abs(np.asarray(cfg[k]) - b[k]).max() <= 5e-7, k
    assert np.array_equal(b["type"], cfg["type"]) and np.array_equal(b["id"], cfg["id"])
    assert b["groups"] == {"all": 1}
    ((ftype, handle, elems),) = b["fixes"]
    assert (ftype, handle) == ("Rigid", "rigid")
    aim = elems[0]
    assert aim.tag == "atomsInMolecule" and int(aim.get("n")) == len(box["mol"])
    ids = np.array(aim.text.split(), int).reshape(-1, 4)
    assert np.array_equal(ids, box["mol"])
    assert elems[1].get("type") == "TIP4P/2005" and elems[2].get("fourSite") == "true"
