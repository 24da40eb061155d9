"""Text output in the reference's formats.

Two families, as the reference has them:

- the C program's files (output.c:104-560): `<prefix>corrfunc.dat`,
  `<prefix>multipoles.dat`, `<prefix>covariance_multipoles.dat`,
  `<prefix>covariance_average_multipoles.dat`, `<prefix>background.dat` and
  `<prefix>integral<i>.dat`, every number written "%.10e";
- the Python command line's tables (python/coffe/__main__.py:52-160), numpy.savetxt
  with a tab delimiter and one header line; `python -m coffe_amd` writes them.

Records are the objects of coffe_amd.representation; separations are in Mpc, as the
records hold them.
"""
import os

import numpy as np

# the contribution names of the header line, in the reference's order (output.c:79-100)
_CONTRIBUTIONS = (("den", "has_density"), ("rsd", "has_rsd"), ("d1", "has_d1"), ("d2", "has_d2"),
                  ("g1", "has_g1"), ("g2", "has_g2"), ("g3", "has_g3"), ("g4", "has_g4"), ("g5", "has_g5"),
                  ("len", "has_lensing"))

# background columns: name, the background table it reads, the unit of the header
# (output.c:201-300)
BACKGROUND_COLUMNS = {
    "z": ("z", ""), "a": ("a", ""), "H": ("Hz", "[1/Mpc]"), "conformal_H": ("conformal_Hz", "[1/Mpc]"),
    "conformal_H_prime": ("conformal_Hz_prime", "[1/Mpc^2]"), "D1": ("D1", ""), "f": ("f", ""),
    "comoving_distance": ("chi", "[Mpc]"),
}
DEFAULT_BACKGROUND = ("z", "a", "H", "conformal_H", "conformal_H_prime", "D1", "f", "comoving_distance")

# the integral tables in the order of their file numbers (output.c:407-416)
INTEGRAL_FILES = ((0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1))


def contributions_header(flags):
    """'# correlation contributions: den rsd ... \\n' for a dict of has_* flags."""
    names = "".join(f"{short} " for short, key in _CONTRIBUTIONS if flags.get(key, False))
    return f"# correlation contributions: {names}\n"


def _append(path, header, lines):
    # the C writers open these in append mode (output.c:120, 152, 175)
    with open(path, "a", encoding="utf-8") as fh:
        fh.write(header)
        fh.writelines(lines)


def write_corrfunc(path, records, flags):
    """output_corrfunc (output.c:104-133): 'z_mean sep mu value' per line."""
    _append(path, contributions_header(flags) + "# z_mean, sep, mu, value\n",
            [f"{r.z:.10e} {r.r:.10e} {r.mu:.10e} {r.value:.10e}\n" for r in records])


def write_multipoles(path, records, flags):
    """output_multipoles (output.c:136-165): 'z_mean l sep value' per line."""
    _append(path, contributions_header(flags) + "# z_mean, l, sep, value\n",
            [f"{r.z:.10e} {int(r.l)} {r.r:.10e} {r.value:.10e}\n" for r in records])


def write_covariance(path, records, flags):
    """output_covariance (output.c:168-199): 'z_mean l1 l2 sep1 sep2 value' per line."""
    _append(path, contributions_header(flags) + "# z_mean, l1, l2, sep1, sep2, value\n",
            [f"{r.z:.10e} {int(r.l1)} {int(r.l2)} {r.r1:.10e} {r.r2:.10e} {r.value:.10e}\n" for r in records])


def write_average_covariance(path, records):
    """covariance_average_multipoles.dat (output.c:383-402), written afresh:
    'sep1 sep2 l1 l2 z_min z_max value' per line under the reference's header."""
    with open(path, "w", encoding="utf-8") as fh:
        fh.write("# sep[Mpc/h]\tsep[Mpc/h]\tresult\n")
        fh.writelines(f"{r.r1:.10e} {r.r2:.10e} {int(r.l1)} {int(r.l2)} {r.z_min:.10e} {r.z_max:.10e} "
                      f"{r.value:.10e}\n" for r in records)


def write_columns(path, header, sep, columns):
    """write_ncol_null (common.c:368-412): one row per sample, every value "%.10e"
    followed by the separator, also the last one."""
    columns = [np.asarray(c, dtype=np.float64) for c in columns]
    with open(path, "w", encoding="utf-8") as fh:
        if header is not None:
            fh.write(header)
        for row in zip(*columns):
            fh.write("".join(f"{v:.10e}{sep}" for v in row) + "\n")


def write_background(path, background, columns=DEFAULT_BACKGROUND, sep="\t"):
    """output_background (output.c:201-322): the chosen columns of the background
    tables (a dict as coffe_amd.background.compute returns) at their samples, headed
    '# 1)z<sep>2)a<sep>3)H[1/Mpc]<sep>...'."""
    unknown = [c for c in columns if c not in BACKGROUND_COLUMNS]
    if unknown:
        raise ValueError(f"unknown background columns {unknown}; choose from {list(BACKGROUND_COLUMNS)}")
    header = "# " + "".join(f"{i + 1}){c}{BACKGROUND_COLUMNS[c][1]}{sep}" for i, c in enumerate(columns)) + "\n"
    write_columns(path, header, sep, [background[BACKGROUND_COLUMNS[c][0]] for c in columns])


def write_integrals(prefix, tables, divergent=None, renormalization=None):
    """The integral tables (output.c:404-560): `<prefix>integral<i>.dat` for the
    eight (n, l) of INTEGRAL_FILES from tables[(n, l)] = (r, values), headed
    '# n = <n>, l = <l>'; with divergent = (r, values) also `<prefix>integral8.dat`,
    and with renormalization = (chi, grid) `<prefix>integral8_renormalization.dat`,
    one 'chi_i chi_j value' line ("%e") per pair, grid[j, i] the value at (chi_i, chi_j)."""
    for i, (n, l) in enumerate(INTEGRAL_FILES):
        x, y = tables[(n, l)]
        write_columns(f"{prefix}integral{i}.dat", f"# n = {n}, l = {l}\n", " ", [x, y])
    if divergent is not None:
        write_columns(f"{prefix}integral8.dat", "# n = 4, l = 0\n", " ", list(divergent))
    if renormalization is not None:
        chi, grid = renormalization
        with open(f"{prefix}integral8_renormalization.dat", "w", encoding="utf-8") as fh:
            fh.writelines(f"{chi[i]:e} {chi[j]:e} {grid[j, i]:e}\n"
                          for i in range(len(chi)) for j in range(len(chi)))


def output_prefix(output_path, output_prefix="", timestamp=None):
    """The path prefix of the C program's files (output.c:336-350): the directory is
    created if missing; the prefix "$TIME" becomes '<timestamp>_'."""
    os.makedirs(output_path, exist_ok=True)
    if not output_path.endswith("/"):
        output_path += "/"
    if output_prefix == "$TIME":
        import time
        stamp = timestamp or time.strftime("%Y-%m-%d-%H-%M-%S")
        return f"{output_path}{stamp}_"
    return f"{output_path}{output_prefix}"


# ---- the Python command line's tables (python/coffe/__main__.py:52-160) --------------
_SAVETXT = dict(delimiter="\t", encoding="utf-8")


def save_background_table(path, cosmo):
    zs = np.linspace(0, 15, cosmo.background_sampling)
    cols = [zs] + [cosmo._bg_column(key, zs) for key in ("a", "Hz", "conformal_Hz", "conformal_Hz_prime", "D1", "f",
                                                          "chi")]
    np.savetxt(path, np.transpose(cols), fmt="%.10e", header="\t".join(
        ["1)z", "2)a", "3)H[1/Mpc]", "4)conformal_H[1/Mpc]", "5)conformal_H_prime[1/Mpc^2]", "6)D1", "7)f",
         "8)comoving_distance[Mpc]"]), **_SAVETXT)


def save_corrfunc_table(path, records):
    np.savetxt(path, np.array([[r.r, r.mu, r.z, r.value] for r in records]), fmt="%.10e",
               header="\t".join(["r[Mpc]", "mu", "z", "corrfunc_value"]), **_SAVETXT)


def save_multipoles_table(path, records):
    np.savetxt(path, np.array([[r.r, r.l, r.z, r.value] for r in records]), fmt=["%.10e", "%d", "%.10e", "%.10e"],
               header="\t".join(["r[Mpc]", "l", "z", "multipoles_value"]), **_SAVETXT)


def save_average_multipoles_table(path, records):
    np.savetxt(path, np.array([[r.r, r.l, r.z_min, r.z_max, r.value] for r in records]),
               fmt=["%.10e", "%d", "%.10e", "%.10e", "%.10e"],
               header="\t".join(["r[Mpc]", "l", "z_min", "z_max", "avg_multipoles_value"]), **_SAVETXT)


def save_covariance_table(path, records):
    np.savetxt(path, np.array([[r.r1, r.r2, r.l1, r.l2, r.z, r.value] for r in records]),
               fmt=["%.10e", "%.10e", "%d", "%d", "%.10e", "%.10e"],
               header="\t".join(["r1[Mpc]", "r2[Mpc]", "l1", "l2", "z", "covariance_multipoles_value"]), **_SAVETXT)


def save_average_covariance_table(path, records):
    np.savetxt(path, np.array([[r.r1, r.r2, r.l1, r.l2, r.z_min, r.z_max, r.value] for r in records]),
               fmt=["%.10e", "%.10e", "%d", "%d", "%.10e", "%.10e", "%.10e"],
               header="\t".join(["r1[Mpc]", "r2[Mpc]", "l1", "l2", "z_min", "z_max",
                                 "covariance_avg_multipoles_value"]), **_SAVETXT)
