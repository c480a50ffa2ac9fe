"""JSON-lines metrics records: one JSON object per line, appended to a file
per run (the bench's result line, or any caller's), each stamped with the
UTC time and the host, so records of many runs collect in one file and read
back with read_records.  The reference logs through an ExpLogger and plots
(logger.py:90-146) and keeps no machine-readable record."""
from __future__ import annotations

import datetime
import json
import os
import socket


def append_record(path, record: dict, **meta) -> dict:
    """Append `record` (plus `meta` and a time / host stamp) as one JSON line
    to `path` (created with its folder if missing); returns what was written."""
    rec = {"time": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
           "host": socket.gethostname(), **meta, **record}
    folder = os.path.dirname(os.path.abspath(os.fspath(path)))
    os.makedirs(folder, exist_ok=True)
    line = json.dumps(rec, sort_keys=False)
    with open(path, "a") as f:
        f.write(line + "\n")
    return rec


def read_records(path) -> list:
    """Every record of a JSON-lines file, in order (blank lines skipped)."""
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]
