"""Command-line drop-in for applications/crystalPlasticity/main.cc: python -m plasticity_amd.cli prm.prm"""
import sys

from .model import main as run_prm


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("Provide name of a parameter file.", file=sys.stderr)  # main.cc:24-26
        return 1
    run_prm(argv[0])
    return 0


if __name__ == "__main__":
    sys.exit(main())
