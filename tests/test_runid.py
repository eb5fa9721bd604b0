"""Run identifiers (pfc_hip.io.runid), the reference's io/runid.py contract."""
import datetime as dt
import hashlib

from pfc_hip.io import config_hash, run_id


def test_hash_is_sha256_of_canonical_json():
    cfg = {"b": 1, "a": [1.5, 2], "c": {"y": None, "x": "s"}}
    blob = '{"a":[1.5,2],"b":1,"c":{"x":"s","y":null}}'
    assert config_hash(cfg) == hashlib.sha256(blob.encode()).hexdigest()[:8]
    assert config_hash(dict(reversed(list(cfg.items())))) == config_hash(cfg)
    assert config_hash({**cfg, "b": 2}) != config_hash(cfg)
    assert len(config_hash(cfg, length=12)) == 12


def test_run_id_format():
    rid = run_id("logpfc_extv", {"dt": 0.1}, date=dt.date(2026, 9, 14))
    day, model, h = rid.split("-", 2)
    assert day == "20260914" and model == "logpfc_extv" and h == config_hash({"dt": 0.1})
    # values JSON cannot hold are written by repr
    assert config_hash({"x": complex(1, 2)}) == hashlib.sha256(b'{"x":"(1+2j)"}').hexdigest()[:8]
