"""The command line (python/coffe/__main__.py:1-170): reads a settings file with
Coffe.from_file, computes one kind of output and writes it as a tab-separated table.

    python -m coffe_amd -s settings.cfg -o out.dat -k mp -p pk.dat

The reference takes its power spectrum from CLASS; this engine has none, so the kinds
other than `bg` need -p, a two-column file of k [1/Mpc] and P(k) [Mpc^3] at z = 0.
"""
from argparse import ArgumentParser

import numpy as np

from . import output
from .coffe import Coffe

KINDS = ("bg", "cf", "mp", "cov", "avg_cov", "avg_mp")


def parse_args(argv=None):
    parser = ArgumentParser(prog="COFFE", description="The command line interface to the COFFE engine")
    parser.add_argument("-s", "--settings", type=str, required=True, help="The settings file to use")
    parser.add_argument("-o", "--output", type=str, required=True, help="In which file to store the output")
    parser.add_argument("-k", "--kind", type=str, required=True, choices=KINDS, help="The kind of output to compute.")
    parser.add_argument("-p", "--power-spectrum", type=str, default=None,
                        help="Two columns: k [1/Mpc] and the linear P(k) [Mpc^3] at z = 0")
    parser.add_argument("-v", "--version", action="version", version="%(prog)s 3.0.0")
    return parser.parse_args(argv)


def compute_output(args):
    cosmo = Coffe.from_file(args.settings)
    if args.kind == "bg":
        output.save_background_table(args.output, cosmo)
        return
    if args.power_spectrum is None:
        raise SystemExit(f"-k {args.kind} needs a power spectrum (-p FILE): there is no CLASS in this engine")
    k, pk = np.loadtxt(args.power_spectrum, unpack=True)[:2]
    cosmo.set_power_spectrum_linear(k, pk)
    if args.kind == "cf":
        output.save_corrfunc_table(args.output, cosmo.compute_corrfunc_bulk())
    elif args.kind == "mp":
        output.save_multipoles_table(args.output, cosmo.compute_multipoles_bulk())
    elif args.kind == "avg_mp":
        output.save_average_multipoles_table(args.output, cosmo.compute_average_multipoles_bulk())
    elif args.kind == "cov":
        output.save_covariance_table(args.output, cosmo.compute_covariance_bulk())
    else:
        output.save_average_covariance_table(args.output, cosmo.compute_average_covariance_bulk())


def main(argv=None):
    compute_output(parse_args(argv))


if __name__ == "__main__":
    main()
