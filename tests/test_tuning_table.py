"""The committed autotuner table parses and names only variants the engine accepts."""
import json
import os

from conftest import PKG

TABLE = os.path.join(PKG, "tuning", "gnss_tuning.json")


def test_table_is_valid():
    t = json.load(open(TABLE))
    assert t, "empty tuning table"
    for key, v in t.items():
        prec, nx, ny, nz, s = key.split()
        s = s[:-1] if s.endswith("b") else s  # "b": entry of a batched (ensemble) context of the grid
        assert prec in ("f32", "f64") and int(s.rstrip("c")) >= (2 if s.endswith("c") else 0)
        assert {"cx", "cy", "zf", "xf", "cp"} <= set(v) <= {"cx", "cy", "zf", "xf", "cp", "xc", "zp", "kp", "cb", "lp", "yc"}
        assert v.get("cb", 0) in (0, 1) and v.get("lp", 1) in (0, 1, 2, 3)
        assert v.get("kp", 0) in range(5)
        allowed = {4, 8} if prec == "f64" else {8, 16}
        assert v["cx"] in allowed and v["cy"] in allowed and v.get("xc", v["cx"]) in allowed
        assert v.get("yc", 0) in allowed | {0}
        # x modes 4 and 5 (late derivatives) are not built for the colloid
        assert v["xf"] in ((0, 1, 2, 3) if s.endswith("c") else (0, 1, 2, 3, 4, 5))
        assert v["cp"] in (0, 1) and v.get("zp", 0) in (0, 1)
        # fused z variant 9 (two rows per transform) is not built for the colloid nor beyond nz 512
        pair = () if s.endswith("c") or int(nz) > 512 else (9,)
        assert v["zf"] in ((0, 3, 6, 7, 8) if prec == "f64" else (0, 1, 2, 3, 4, 5, 6, 7)) + pair
    # one entry per line: the engine scans the file line by line for its key
    lines = [l for l in open(TABLE).read().splitlines() if l.startswith('"')]
    assert len(lines) == len(t)


def test_measured_workloads_are_tuned():
    t = json.load(open(TABLE))
    for key in ["f32 512 512 512 1", "f64 512 512 512 1", "f64 256 256 256 2", "f32 1024 1024 1024 1",
                "f32 64 64 64 1", "f32 96 96 96 1", "f32 128 128 128 1", "f32 512 512 512 2c",
                "f64 512 512 512 2c"]:
        assert key in t, key


def test_autotuner_merges_into_the_table(tmp_path):
    """Re-tuning replaces only the keys it tunes: every other entry (the batched "<key>b" ones
    included) survives, and the untuned row padding kp keeps its table value."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("autotune", os.path.join(PKG, "autotune.py"))
    at = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(at)
    path = str(tmp_path / "t.json")
    old = json.load(open(TABLE))
    at.write_table(path, old)
    table = at.read_table(path)
    assert table == old
    key = "f32 512 512 512 1"
    table[key] = dict(table[key], kp=2)
    rep = {"key": key, "choice": "cx=8 cy=16 zf=6 xf=4 cp=1 xc=16 zp=1 kp=0 cb=1 lp=1 yc=0"}
    table[key] = at.choice_of(rep, table[key])
    at.write_table(path, table)
    new = json.load(open(path))
    assert new[key]["cx"] == 8 and new[key]["zf"] == 6 and new[key]["kp"] == 2
    for k in old:
        if k != key:
            assert new[k] == old[k], k
    assert any(k.endswith("b") for k in new)
