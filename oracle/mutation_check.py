"""Oracle mutation check (test infrastructure, CPU only).

Plants one-line mistakes (a dropped term, a flipped stagger, an off-by-one
bound, a wrong grid step) in a copy of the oracle under /tmp/mut, rebuilds
it, and runs the CPU oracle pins against each copy: every mutation must make
at least one pin fail.  Usage: python oracle/mutation_check.py [substring]
(the optional substring keeps the mutations whose description contains it).
"""
import os, shutil, subprocess, sys
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MUTS = [
 ("hpp", "ttxx += (R)C8[i] * ((L[p1] + 2 * M[p1]) * f.txx[p1] - (L[p0] + 2 * M[p0]) * f.txx[p0]);",
         "ttxx += (R)C8[i] * ((L[p1] + M[p1]) * f.txx[p1] - (L[p0] + 2 * M[p0]) * f.txx[p0]);", "adjoint_v: dropped 2mu"),
 ("hpp", "ttxz += (R)C8[i] * (M[c + i - 1] * f.txz[c + i - 1] - M[c - i] * f.txz[c - i]);",
         "ttxz += (R)C8[i] * (M[c + i] * f.txz[c + i] - M[c - i + 1] * f.txz[c - i + 1]);", "adjoint_v: flipped txz stagger"),
 ("hpp", "R nvx_x = (d1 * f.vx_x[c] + dir * br * (dt / dx) * ttxx + dir * br * (dt / dx) * ttzz) / d2;",
         "R nvx_x = (d1 * f.vx_x[c] + dir * br * (dt / dx) * ttxx) / d2;", "adjoint_v: dropped lambda*tzz term"),
 ("hpp", "            f.txx[c] = f.txx_x[c];\n", "            f.txx[c] = f.txx_x[c] + f.txx_z[c];\n", "adjoint_s: txx sums both splits"),
 ("hpp", "            f.txz_x[c] = (d1 * f.txz_x[c] + dir * (dt / dx) * dvzx) / d2;",
         "            f.txz_x[c] = (d1 * f.txz_x[c] + dir * (dt / dz) * dvzx) / d2;", "adjoint_s: dz for dx"),
 ("hpp", "txx_x += (R)C8[i] * (CP(xp) * s.txx[xp] - CP(xm) * s.txx[xm]);",
         "txx_x += (R)C8[i] * (CX(xp) * s.txx[xp] - CX(xm) * s.txx[xm]);", "born adjoint: wrong modulus"),
 ("hpp", "txz_z += (R)C8[i] * (CS(c + i - 1) * s.txz[c + i - 1] - CS(c - i) * s.txz[c - i]);",
         "txz_z += (R)C8[i] * (CS(c + i) * s.txz[c + i] - CS(c - i + 1) * s.txz[c - i + 1]);", "born adjoint: txz_z stagger"),
 ("hpp", "bvx[c] = bvx[c] + dir * br * (txx_x / dx + tzz_x / dx + txz_z / dz);",
         "bvx[c] = bvx[c] + dir * br * (txx_x / dx + txz_z / dz);", "born adjoint: dropped tzz_x"),
 ("hpp", "            s.gx_vx[c] = dvx / dx;", "            s.gx_vx[c] = dvx / dz;", "recon strain: dz for dx"),
 ("hpp", "            s.gz_vx[c] = dvxz / dz;", "            s.gz_vx[c] = dvzx / dx;", "recon strain: swapped shear strains"),
 ("hpp", "if (ix >= pml && ix < nxp - pml && iz >= pml - N && iz < pml + N)",
         "if (ix >= pml && ix < nxp - pml && iz >= pml - N && iz < pml + N - 1)", "edge_io: off-by-one band"),
 ("hpp", "io(e.left, (size_t)it * 2 * N * nz + (size_t)(ix - (pml - N)) * nz + (iz - pml));",
         "io(e.left, (size_t)it * 2 * N * nz + (size_t)(ix - (pml - N)) * nz + (iz - pml + 1));", "edge_io: shifted left band"),
 ("hpp", "R mg = -2 * (s.gx_vx[c] * bk.tzz[c] + s.gz_vz[c] * bk.txx[c]) + (s.gz_vx[c] + s.gx_vz[c]) * bk.txz[c];",
         "R mg = -2 * (s.gx_vx[c] * bk.txx[c] + s.gz_vz[c] * bk.tzz[c]) + (s.gz_vx[c] + s.gx_vz[c]) * bk.txz[c];", "image: PS pairs swapped"),
 ("hpp", "size_t o = (size_t)(col0 + ix) * allnz + iz;\n            pp[o] += cpp;",
         "size_t o = (size_t)(col0 + ix + 1) * allnz + iz;\n            pp[o] += cpp;", "image: column off by one"),
 ("cpp", "tvz += (R)C8[i] * (vz[c + i] - vz[c - i + 1]);", "tvz += (R)C8[i] * (vz[c + i - 1] - vz[c - i]);", "acoustic ref: p z stagger"),
 ("cpp", "vz[c] = (d1 * vz[c] - ((R)1 / rr[c]) * (dt / dx) * tp) / d2;", "vz[c] = (d1 * vz[c] - ((R)1 / rr[c]) * (dt / dz) * tp) / d2;", "acoustic ref: vz dt/dz (the reference uses dt/dx)"),
 ("cpp", "    for (int ix = NOP - 1; ix <= nxp - NOP - 1; ++ix)\n        for (int iz = NOP - 1; iz <= nzp - NOP - 1; ++iz) {\n            size_t c = g.at(ix, iz);\n            R d1 = 1 - dt * damp[c] / 2, d2 = 1 + dt * damp[c] / 2;\n            R tvx",
         "    for (int ix = NOP; ix <= nxp - NOP - 1; ++ix)\n        for (int iz = NOP - 1; iz <= nzp - NOP - 1; ++iz) {\n            size_t c = g.at(ix, iz);\n            R d1 = 1 - dt * damp[c] / 2, d2 = 1 + dt * damp[c] / 2;\n            R tvx", "acoustic ref: p lower bound"),
 ("rot", "            f.vsx[c] = (d1 * f.vsx[c] + dir * s2 * (dt / dz) * oz) / d2;",
         "            f.vsx[c] = (d1 * f.vsx[c] + dir * s2 * (dt / dx) * oz) / d2;", "rot_v: dx for dz in vsx"),
 ("rot", "tx += (R)C8[i] * (f.th[c + (size_t)(i - 1) * nzp] - f.th[c - (size_t)i * nzp]);",
         "tx += (R)C8[i] * (f.th[c + (size_t)i * nzp] - f.th[c - (size_t)(i - 1) * nzp]);", "rot_v: flipped theta x stagger"),
 ("rot", "            f.vpz[c] = (d3 * f.vpz[c] + dir * p2 * (dt / dz) * tz) / d4;",
         "            f.vpz[c] = (d3 * f.vpz[c] + dir * s2 * (dt / dz) * tz) / d4;", "rot_v: vs^2 for vp^2"),
 ("rot", "            f.om_x[c] = (d1 * f.om_x[c] - dir * (dt / dx) * zx) / d2;",
         "            f.om_x[c] = (d1 * f.om_x[c] + dir * (dt / dx) * zx) / d2;", "rot_p: omega_x sign"),
 ("rot", "            VS[c] = xz / g.dz - zx / g.dx;", "            VS[c] = xz / g.dz + zx / g.dx;", "helmholtz: curl sign"),
 ("rot", "                xz += (R)C8[i] * (vx[c + i - 1] - vx[c - i]);",
         "                xz += (R)C8[i] * (vx[c + i] - vx[c - i + 1]);", "helmholtz: flipped vx z stagger"),
 ("rot", "    if (sxz) { sxz[0] = -sz / g.dz; sxz[1] = sx / g.dx; }",
         "    if (sxz) { sxz[0] = sz / g.dz; sxz[1] = sx / g.dx; }", "vector: S sign"),
 ("rot", "    const bool dn = p >= 0;", "    const bool dn = p <= 0;", "poynting: direction flipped"),
 ("rot", "    const R cc = (fu[0] * bd[0] + fu[1] * bd[1]) + (fd[0] * bu[0] + fd[1] * bu[1]);",
         "    const R cc = (fu[0] * bu[0] + fu[1] * bu[1]) + (fd[0] * bd[0] + fd[1] * bd[1]);", "poynting image: same-direction pairs"),
 ("cpp", "        rot_p(S.g, f, m.dampx.data(), m.dampz.data(), -1);\n        f.th[src] += w;\n        rot_v(S.g, f, m.vp.data(), m.vs.data(), m.dampx.data(), m.dampz.data(), -1);",
         "        rot_p(S.g, f, m.dampx.data(), m.dampz.data(), -1);\n        rot_v(S.g, f, m.vp.data(), m.vs.data(), m.dampx.data(), m.dampz.data(), -1);\n        f.th[src] += w;",
         "rot reversal: the reference's injection order"),
]
FILES = {"hpp": "ds2d_oracle.hpp", "cpp": "ds2d_oracle.cpp", "rot": "ds2d_oracle_rot.hpp"}
TESTS = ["tests/test_oracle_kernel_pins.py", "tests/test_oracle_pins.py", "tests/test_oracle_physics.py",
         "tests/test_oracle_eborn.py", "tests/test_oracle_rot.py"]
root = "/tmp/mut"
res = []
only = sys.argv[1] if len(sys.argv) > 1 else None  # optional substring filter on the descriptions
for k, (which, old, new, desc) in enumerate(MUTS):
    if only and only not in desc:
        continue
    d = f"{root}/{k}"
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    for sub in ("tests", "oracle", "ds2d_amd"):
        shutil.copytree(f"{REPO}/{sub}", f"{d}/{sub}", ignore=shutil.ignore_patterns("_build", "bin", "*.so", "__pycache__", "_ref"))
    shutil.copy(f"{REPO}/pytest.ini", d)
    f = f"{d}/oracle/{FILES[which]}"
    src = open(f).read()
    assert src.count(old) == 1, (desc, src.count(old))
    open(f, "w").write(src.replace(old, new))
    r = subprocess.run(["make", "-s", "-C", f"{d}/oracle", f"{d}/oracle/libds2d_oracle.so"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *TESTS], cwd=d, capture_output=True, text=True)
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    res.append((desc, r.returncode != 0, failed[:1]))
    print(f"{'CAUGHT' if r.returncode else 'MISSED'}  {desc}  {failed[:1]}", flush=True)
print(f"{sum(c for _, c, _ in res)}/{len(res)} mutations caught")
