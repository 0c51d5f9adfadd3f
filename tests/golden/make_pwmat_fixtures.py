"""Copies the PWmat / UPF data files the reference's own tests read into
tests/golden/pwmat/ (byte copies; every k-point of every OUT*.GKK).  Run in
a checkout that has the reference at REF (default /root/reference); GPU runs
only see the copies.

Sources (reference test data): test/data_nonscf/{OUT1.GKK (7 k-points on 2
nodes), OUT2.GKK, IN.VR, OUT1.EIGEN, OUT2.EIGEN, atom.config,
Si.SG15.PBE.UPF}, test/data_scf/{OUT.GKK, OUT.VR, OUT.EIGEN, atom.config,
Al.SG15.PBE.UPF, N.SG15.PBE.UPF}, test/data_upf/Ge-spd-high.PD04.PBE.UPF
(norm-conserving, projectors with l = 0, 1, 2).
"""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pwmat")


def main():
    os.makedirs(OUT, exist_ok=True)
    t = os.path.join(REF, "test")
    for src, dst in [("data_nonscf/OUT1.GKK", "si8_OUT1.GKK"), ("data_nonscf/OUT2.GKK", "si8_OUT2.GKK"),
                     ("data_nonscf/IN.VR", "si8_IN.VR"), ("data_nonscf/OUT1.EIGEN", "si8_OUT1.EIGEN"),
                     ("data_nonscf/OUT2.EIGEN", "si8_OUT2.EIGEN"),
                     ("data_nonscf/atom.config", "si8_atom.config"),
                     ("data_nonscf/Si.SG15.PBE.UPF", "Si.SG15.PBE.UPF"),
                     ("data_scf/OUT.GKK", "aln_OUT.GKK"),
                     ("data_scf/OUT.VR", "aln_OUT.VR"), ("data_scf/OUT.EIGEN", "aln_OUT.EIGEN"),
                     ("data_scf/atom.config", "aln_atom.config"),
                     ("data_scf/Al.SG15.PBE.UPF", "Al.SG15.PBE.UPF"),
                     ("data_scf/N.SG15.PBE.UPF", "N.SG15.PBE.UPF"),
                     ("data_upf/Ge-spd-high.PD04.PBE.UPF", "Ge-spd-high.PD04.PBE.UPF")]:
        shutil.copyfile(os.path.join(t, src), os.path.join(OUT, dst))


if __name__ == "__main__":
    main()
