#!/usr/bin/env python3
"""Regenerate the golden fixtures from the reference itself.

Builds oracle/_ref/ref_harness (our driver compiled against the ViennaEMC
headers where they lie, see oracle/Makefile), runs it in a scratch folder
under oracle/_ref/, and packs what it recorded, together with the result
files the reference's emcSimulation wrote there (text, as bytes), into
tests/golden/reference_golden.npz.  Only runs where the reference sources
are present; the committed .npz is what the tests read.

Usage: python tests/golden/make_golden.py [--ref /root/reference]
"""
import argparse
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(ROOT, "oracle")

# trajectory snapshots kept for k, E and tau (every SNAP-th step and the last)
SNAP = 5


def parse(path):
    out = {}
    with open(path) as f:
        while True:
            head = f.readline()
            if not head:
                break
            name, dt, n = head[1:].split()
            line = f.readline().split()
            assert len(line) == int(n), name
            if dt == "f8":
                out[name] = np.array([float(x) for x in line], dtype=np.float64)
            elif dt == "u8":
                out[name] = np.array([int(x) for x in line], dtype=np.uint64)
            else:
                out[name] = np.array([int(x) for x in line], dtype=np.int64)
    return out


def shrink(d):
    """Subsample per-step trajectory arrays and narrow integer types."""
    res = {}
    for name, a in d.items():
        parts = name.split(".")
        if (len(parts) == 3 and parts[1] == "traj" and parts[2] in ("k", "E", "tau", "pos")
                and not (parts[0] == "device" and parts[2] == "pos")):
            meta = d[parts[0] + ".traj.meta"]
            np_, ns = int(meta[0]), int(meta[1])
            per = 3 if parts[2] in ("k", "pos") else 1
            a = a.reshape(ns, np_, per)
            keep = sorted(set(list(range(SNAP - 1, ns, SNAP)) + [ns - 1]))
            res[name] = a[keep].reshape(-1)
            res[parts[0] + ".traj.snap_steps"] = np.array(keep, dtype=np.int32)
            continue
        if a.dtype == np.int64 and name.split(".")[-1] not in ("meta",):
            if a.size and a.min() >= -32768 and a.max() < 32768:
                a = a.astype(np.int16)
            else:
                a = a.astype(np.int32)
        res[name] = a
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=os.path.join(HERE, "reference_golden.npz"))
    args = ap.parse_args()
    if not os.path.isdir(os.path.join(args.ref, "include")):
        print("reference sources not found at %s" % args.ref, file=sys.stderr)
        return 2
    subprocess.check_call(["make", "-C", ORACLE, "ref", "REF=%s" % args.ref])
    work = os.path.join(ORACLE, "_ref", "work")
    os.makedirs(work, exist_ok=True)
    raw = os.path.join(ORACLE, "_ref", "golden_raw.txt")
    with open(os.path.join(ORACLE, "_ref", "harness_stdout.txt"), "w") as log:
        subprocess.check_call([os.path.join(ORACLE, "_ref", "ref_harness"), raw], cwd=work, stdout=log)
    d = shrink(parse(raw))
    # result files written by the reference's emcSimulation::execute (text)
    for name in sorted(os.listdir(work)):
        if name.startswith("sim_") and name.endswith(".txt"):
            with open(os.path.join(work, name), "rb") as f:
                d["simfile." + name[:-4]] = np.frombuffer(f.read(), dtype=np.uint8)
    # scatter-table dumps (initScatterTables -> writeTablesToFiles) of two
    # particle types, written into their own folders
    for sub in ("tab_si_vasileska", "tab_device"):
        folder = os.path.join(work, sub)
        for name in sorted(os.listdir(folder)):
            with open(os.path.join(folder, name), "rb") as f:
                d["tabfile.%s.%s" % (sub[4:], name[:-4])] = np.frombuffer(f.read(), dtype=np.uint8)
    # the bulk example's files of the real handler (particle dumps after
    # initialisation and at the end, the three time series)
    for sub in sorted(os.listdir(work)):
        if sub.startswith("bulk_") and os.path.isdir(os.path.join(work, sub)):
            for name in sorted(os.listdir(os.path.join(work, sub))):
                with open(os.path.join(work, sub, name), "rb") as f:
                    d["bulkfile.%s.%s" % (sub[5:], name[5:-4])] = np.frombuffer(f.read(),
                                                                           dtype=np.uint8)
    d["_meta.compiler"] = np.frombuffer(
        subprocess.check_output(["g++", "--version"]).splitlines()[0], dtype=np.uint8)
    np.savez_compressed(args.out, **d)
    print("wrote %s (%d arrays, %.1f KB)" % (args.out, len(d), os.path.getsize(args.out) / 1024))
    return 0


if __name__ == "__main__":
    sys.exit(main())
