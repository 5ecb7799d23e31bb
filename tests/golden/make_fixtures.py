"""Extract the reference-held golden data used by the oracle tests into tests/golden/.

Run in a container that has the reference checkout (read-only) at /root/reference:

    python tests/golden/make_fixtures.py

Copies (data files only, unchanged or line subsets):
  * samples/NCPA_canonical_profile_zuvwtdp.dat                         -> canonical_profile.dat
  * docs/.../modess_gp_files/code_outputs/tloss_1d.nm, tloss_1d.lossless.nm
        (Modess, canonical profile, 0.1 Hz, az 90, --write_2d_tloss; pre-2.0.1 scaling)
  * docs/.../wmod_gp_files/code_outputs/wtloss_1d.nm, wtloss_1d.lossless.nm
        (WMod, canonical profile, 0.1 Hz, az 90, --write_2d_tloss; pre-2.0.1 scaling)
  * samples/disp_256freqs_0p5hz_NCPA_canonic.dat, lines for 6 frequencies -> disp256_subset.dat
  * docs/.../modbb_gp_files/code_outputs/myDispersionFile.dat, 4 lines -> disp250_subset.dat
  * docs/.../modbb_gp_files/code_outputs/source_spectrum.dat            -> source_spectrum.dat
  * docs/.../modbb_gp_files/code_outputs/myDispersionFile.dat (all 250 lines, parsed, no
        arithmetic) -> disp250.npz: f[250], counts[250], q[total, 4] = Re k, Im k, psi_S, psi_R,
        rho_s, rho_r
  * docs/.../modbb_gp_files/code_outputs/mywavf.dat rows with 880 <= t <= 960 s
        (ModBB pulse1 at 270 km, max celerity 320 m/s, NFFT 16384)       -> mywavf_880_960.dat
"""
import os
import shutil

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
MODESS = os.path.join(REF, "docs/ncpaprop_run_manual_TeX/figs/modess_gp_files/code_outputs")
MODBB = os.path.join(REF, "docs/ncpaprop_run_manual_TeX/figs/modbb_gp_files/code_outputs")
WMOD = os.path.join(REF, "docs/ncpaprop_run_manual_TeX/figs/wmod_gp_files/code_outputs")


def subset(src, dst, line_numbers):
    with open(src) as fh:
        lines = fh.readlines()
    with open(dst, "w") as fh:
        for n in line_numbers:
            fh.write(lines[n])


def main():
    shutil.copyfile(os.path.join(REF, "samples/NCPA_canonical_profile_zuvwtdp.dat"),
                    os.path.join(HERE, "canonical_profile.dat"))
    for f in ("tloss_1d.nm", "tloss_1d.lossless.nm"):
        shutil.copyfile(os.path.join(MODESS, f), os.path.join(HERE, f))
    for f in ("wtloss_1d.nm", "wtloss_1d.lossless.nm"):
        shutil.copyfile(os.path.join(WMOD, f), os.path.join(HERE, f))
    # 0-based line indices: f = (i+1)/512 Hz
    subset(os.path.join(REF, "samples/disp_256freqs_0p5hz_NCPA_canonic.dat"),
           os.path.join(HERE, "disp256_subset.dat"), [0, 4, 25, 64, 127, 255])
    subset(os.path.join(MODBB, "myDispersionFile.dat"), os.path.join(HERE, "disp250_subset.dat"),
           [0, 4, 62, 249])
    shutil.copyfile(os.path.join(MODBB, "source_spectrum.dat"),
                    os.path.join(HERE, "source_spectrum.dat"))
    f, counts, rows, rho = [], [], [], None
    with open(os.path.join(MODBB, "myDispersionFile.dat")) as fh:
        for line in fh:
            v = line.split()
            n = int(v[1])
            f.append(float(v[0]))
            counts.append(n)
            rho = (float(v[2]), float(v[3]))
            rows.append(np.array(v[4:4 + 4 * n], dtype=float).reshape(n, 4))
    np.savez_compressed(os.path.join(HERE, "disp250.npz"), f=np.array(f), counts=np.array(counts),
                        q=np.concatenate(rows), rho_s=rho[0], rho_r=rho[1])
    with open(os.path.join(MODBB, "mywavf.dat")) as fh, \
            open(os.path.join(HERE, "mywavf_880_960.dat"), "w") as out:
        for line in fh:
            if 880.0 <= float(line.split()[0]) <= 960.0:
                out.write(line)


if __name__ == "__main__":
    main()
