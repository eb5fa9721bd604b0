"""Run identifiers `YYYYMMDD-<model>-<hash8>` (the reference's io/runid.py:15-24).

The hash is the first 8 hex digits of SHA-256 over the config serialised as
compact JSON with sorted keys (values JSON cannot hold written as their
repr), so key order does not change it and any changed value does.  The same
config run on the same day gets the same id; sweeps and served jobs use it
as their result key.
"""
from __future__ import annotations

import datetime as _dt
import hashlib
import json
from typing import Any, Mapping


def config_hash(config: Mapping[str, Any], length: int = 8) -> str:
    text = json.dumps(config, sort_keys=True, separators=(",", ":"), default=repr)
    return hashlib.sha256(text.encode()).hexdigest()[:length]


def run_id(model: str, config: Mapping[str, Any], date: _dt.date | None = None) -> str:
    return f"{(date or _dt.date.today()):%Y%m%d}-{model}-{config_hash(config)}"
