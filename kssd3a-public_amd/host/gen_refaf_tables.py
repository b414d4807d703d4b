#!/usr/bin/env python3
"""Writes refaf_hgb.bin, the parameter tables of the reference's calibrated
ANI model (a 160-tree gradient-boosted regression over 11 features,
model_refaf_hgb.h:7-7566 in yhg926/KSSD3A-public, Apache-2.0), from the
reference header.  The tables are fitted model parameters: calibrated ANI
values only agree with the reference if they are identical, so they are
carried over as data.  Run where the reference tree is present:

    python kssd3a-public_amd/host/gen_refaf_tables.py [path/to/model_refaf_hgb.h]

Binary layout (little endian), read by kssd_report.c:
    char[8]  magic "KSSDHGB1"
    u32      n_features, n_trees, n_nodes, 0
    f64      baseline
    u32      tree_offsets[n_trees + 1], zero padding to a multiple of 8 bytes
    f64      threshold[n_nodes]
    f64      value[n_nodes]
    u16      left[n_nodes], right[n_nodes]
    i8       feature[n_nodes]
    u8       is_leaf[n_nodes], missing_go_left[n_nodes]
"""
import re
import struct
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/model_refaf_hgb.h")


def arrays(text):
    out = {}
    for m in re.finditer(r"static const [a-z ]+ (refaf_hgb_\w+)\[(\d+)\]\s*=\s*\{(.*?)\};", text, re.S):
        name, n, body = m.group(1), int(m.group(2)), m.group(3)
        vals = [v for v in (x.strip() for x in body.split(",")) if v]
        assert len(vals) == n, (name, len(vals), n)
        out[name] = vals
    base = re.search(r"refaf_hgb_baseline\s*=\s*([-0-9.eE+]+);", text).group(1)
    feat = int(re.search(r"#define REFAF_HGB_N_FEATURES (\d+)", text).group(1))
    return out, float(base), feat


def main():
    text = SRC.read_text()
    a, baseline, nfeat = arrays(text)
    offsets = [int(v) for v in a["refaf_hgb_tree_offsets"]]
    n_trees, n_nodes = len(offsets) - 1, offsets[-1]
    blob = bytearray(b"KSSDHGB1")
    blob += struct.pack("<4I", nfeat, n_trees, n_nodes, 0)
    blob += struct.pack("<d", baseline)
    blob += struct.pack(f"<{n_trees + 1}I", *offsets)
    blob += bytes((-len(blob)) % 8)
    blob += struct.pack(f"<{n_nodes}d", *[float(v) for v in a["refaf_hgb_threshold"]])
    blob += struct.pack(f"<{n_nodes}d", *[float(v) for v in a["refaf_hgb_value"]])
    blob += struct.pack(f"<{n_nodes}H", *[int(v) for v in a["refaf_hgb_left"]])
    blob += struct.pack(f"<{n_nodes}H", *[int(v) for v in a["refaf_hgb_right"]])
    blob += struct.pack(f"<{n_nodes}b", *[int(v) for v in a["refaf_hgb_feature_idx"]])
    blob += struct.pack(f"<{n_nodes}B", *[int(v) for v in a["refaf_hgb_is_leaf"]])
    blob += struct.pack(f"<{n_nodes}B", *[int(v) for v in a["refaf_hgb_missing_go_left"]])
    (HERE / "refaf_hgb.bin").write_bytes(bytes(blob))
    print(f"wrote {HERE / 'refaf_hgb.bin'}: {n_trees} trees, {n_nodes} nodes, {len(blob)} bytes")


if __name__ == "__main__":
    main()
