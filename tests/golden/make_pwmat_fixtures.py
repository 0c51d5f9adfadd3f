"""Copies the PWmat / UPF data files the reference's own tests read into
tests/golden/pwmat/, trimming OUT.GKK files to their first k-point (same
Fortran record layout, nkpt = 1).  Run in a checkout that has the reference
at REF (default /root/reference); GPU runs only see the copies.

Sources (reference test data): test/data_nonscf/{OUT2.GKK, IN.VR, OUT2.EIGEN,
atom.config, Si.SG15.PBE.UPF}, test/data_scf/{OUT.GKK, OUT.VR, OUT.EIGEN,
atom.config, Al.SG15.PBE.UPF, N.SG15.PBE.UPF}.
"""
import os
import shutil
import struct
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pwmat")


def records(buf):
    o = 0
    while o < len(buf):
        (n,) = struct.unpack_from("<i", buf, o)
        yield buf[o:o + 8 + n]
        o += 8 + n


def trim_gkk(src, dst, keep=1):
    buf = open(src, "rb").read()
    recs = list(records(buf))
    h = list(struct.unpack("<8i", recs[0][4:-4]))
    nnode, nkpt = h[4], h[5]
    h[5] = keep
    body = struct.pack("<8i", *h)
    out = [struct.pack("<i", len(body)) + body + struct.pack("<i", len(body)), recs[1], recs[2]]
    ng = list(struct.unpack("<%di" % ((len(recs[3]) - 8) // 4), recs[3][4:-4]))
    ng2 = [ng[0]] + ng[1:1 + keep * nnode]
    body = struct.pack("<%di" % len(ng2), *ng2)
    out.append(struct.pack("<i", len(body)) + body + struct.pack("<i", len(body)))
    out += recs[4:4 + 4 * nnode * keep]
    open(dst, "wb").write(b"".join(out))


def main():
    os.makedirs(OUT, exist_ok=True)
    t = os.path.join(REF, "test")
    trim_gkk(os.path.join(t, "data_nonscf", "OUT2.GKK"), os.path.join(OUT, "si8_OUT2_k0.GKK"))
    trim_gkk(os.path.join(t, "data_scf", "OUT.GKK"), os.path.join(OUT, "aln_OUT_k0.GKK"))
    for src, dst in [("data_nonscf/IN.VR", "si8_IN.VR"), ("data_nonscf/OUT2.EIGEN", "si8_OUT2.EIGEN"),
                     ("data_nonscf/atom.config", "si8_atom.config"),
                     ("data_nonscf/Si.SG15.PBE.UPF", "Si.SG15.PBE.UPF"),
                     ("data_scf/OUT.VR", "aln_OUT.VR"), ("data_scf/OUT.EIGEN", "aln_OUT.EIGEN"),
                     ("data_scf/atom.config", "aln_atom.config"),
                     ("data_scf/Al.SG15.PBE.UPF", "Al.SG15.PBE.UPF"),
                     ("data_scf/N.SG15.PBE.UPF", "N.SG15.PBE.UPF")]:
        shutil.copyfile(os.path.join(t, src), os.path.join(OUT, dst))


if __name__ == "__main__":
    main()
