"""Write tests/golden/sac/ from the reference's own m_sac.f90 and m_tar.f90.

Runs oracle/_ref/sac_driver, built by oracle/ref_sac/Makefile from the
unmodified reference sources src/shared/m_std.f90, m_tar.f90 and m_sac.f90.
The driver writes a SAC file with every header field set, one with the
defaults of sac__init, a three-trace CSF file and a two-member tar archive.
Run in the build container, where the reference is mounted:

    make -C oracle/ref_sac && python tests/golden/make_sac_golden.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FILES = ("full.sac", "default.sac", "three.csf", "two.tar")


def run_driver(outdir: str) -> None:
    exe = os.path.join(ROOT, "oracle", "_ref", "sac_driver")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref_sac")], check=True)
    subprocess.run([exe, outdir], check=True)


def main() -> int:
    dst = os.path.join(HERE, "sac")
    os.makedirs(dst, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        run_driver(tmp)
        for fn in FILES:
            shutil.copyfile(os.path.join(tmp, fn), os.path.join(dst, fn))
    print("wrote", ", ".join(FILES))
    return 0


if __name__ == "__main__":
    sys.exit(main())
