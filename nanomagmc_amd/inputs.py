"""Input-file generator for nmc_run (and for the reference's `run`).

Writes a working directory in the reference's layout -- INPUT.dat,
Temps/<name>, Fields/<name>, Js/<name>, Output/ (Old/lib/param_read.cpp:35-151,
Old/lib/field_type.cpp:39-71) -- for the benchmark lattices, including the
2-D interaction tables the reference does not ship:

  2d_nn.dat       : 4 nearest-neighbour bonds, J = 1, D = 0
  2d_nn_neel.dat  : the same bonds with interfacial (Neel) DMI D_ij = z x r_ij
  3d_nn_bloch.dat : 6 bonds, J = 1, bulk (Bloch) DMI D_ij = r_ij
                    (the record order of Old/Js/3d_nn_bloch.dat)

Records are "d ints, J, Dx Dy Dz" as the reference parser reads them.
"""
import os

import numpy as np

JS = {
    "2d_nn.dat": [((0, 1), 1, (0, 0, 0)), ((0, -1), 1, (0, 0, 0)), ((1, 0), 1, (0, 0, 0)), ((-1, 0), 1, (0, 0, 0))],
    # D = z x r : r = +y -> (-1, 0, 0), -y -> (1, 0, 0), +x -> (0, 1, 0), -x -> (0, -1, 0)
    "2d_nn_neel.dat": [((0, 1), 1, (-1, 0, 0)), ((0, -1), 1, (1, 0, 0)), ((1, 0), 1, (0, 1, 0)),
                       ((-1, 0), 1, (0, -1, 0))],
    "3d_nn_bloch.dat": [((0, 0, 1), 1, (0, 0, 1)), ((0, 0, -1), 1, (0, 0, -1)), ((0, 1, 0), 1, (0, 1, 0)),
                        ((0, -1, 0), 1, (0, -1, 0)), ((1, 0, 0), 1, (1, 0, 0)), ((-1, 0, 0), 1, (-1, 0, 0))],
}

# The benchmark lattices as INPUT keys (temperatures listed from Ts[0],
# which the reference uses as its start temperature, main.cpp:46,118).
CONFIGS = {
    "ising2d": dict(keys=dict(ISISING=1, EXCHANGE=1, DMISTREN=0, ISDISTRIB=0, SIZE=64, LATTSHAPE="s",
                              NLATTS=1, BOLTZMANN=1, ISPERIO=1, SAMPSWEEPS=10, EQSWEEPS=1000, NSAMPS=1000,
                              PROTOCOL=2, INTERACTIONS="2d_nn.dat"),
                    temps=list(np.linspace(3.5, 1.5, 21)), fields=[0.0]),
    "heis3d": dict(keys=dict(ISISING=0, EXCHANGE=1, DMISTREN=0, ISDISTRIB=0, SIZE=64, LATTSHAPE="c",
                             NLATTS=1, BOLTZMANN=1, ISPERIO=1, SAMPSWEEPS=5, EQSWEEPS=2000, NSAMPS=2000,
                             PROTOCOL=2, INTERACTIONS="3d_nn_bloch.dat", TRIAL=1, CONEANGLE=1.0),
                   temps=list(np.linspace(2.0, 0.8, 32)), fields=[0.0, 0.05]),
    "skyrmion2d": dict(keys=dict(ISISING=0, EXCHANGE=1, DMISTREN=0.3, ISDISTRIB=0, SIZE=256, LATTSHAPE="s",
                                 NLATTS=1, BOLTZMANN=1, ISPERIO=1, SAMPSWEEPS=10, EQSWEEPS=1000, NSAMPS=500,
                                 PROTOCOL=1, INTERACTIONS="2d_nn_neel.dat"),
                       temps=list(np.linspace(1.0, 0.05, 20)), fields=list(np.linspace(0.0, 0.12, 25))),
    "grains": dict(keys=dict(ISISING=0, EXCHANGE=1, DMISTREN=0.2, ISDISTRIB=1, MEANSIZE=10, SIZEDEV=2.5,
                             LATTSHAPE="x", WEIBULLFACT=8, NLATTS=512, BOLTZMANN=1, ISPERIO=0, SAMPSWEEPS=5,
                             EQSWEEPS=200, NSAMPS=200, PROTOCOL=2, INTERACTIONS="3d_nn_bloch.dat",
                             MINSIZE=4, MAXSIZE=20),
                   temps=list(np.linspace(2.0, 0.1, 16)), fields=[0.0, 0.02, 0.05, 0.1]),
}


def write_js(dirpath, name):
    os.makedirs(os.path.join(dirpath, "Js"), exist_ok=True)
    with open(os.path.join(dirpath, "Js", name), "w") as f:
        for d, J, D in JS[name]:
            f.write(" ".join(str(x) for x in d) + " %g %g %g %g\n" % (J, D[0], D[1], D[2]))


def write_case(dirpath, keys, temps, fields, tname="temps.dat", fname="fields.dat"):
    """A complete working directory for one run."""
    for sub in ("Temps", "Fields", "Js", "Output"):
        os.makedirs(os.path.join(dirpath, sub), exist_ok=True)
    with open(os.path.join(dirpath, "Temps", tname), "w") as f:
        f.write("\n".join("%.9g" % t for t in temps) + "\n")
    with open(os.path.join(dirpath, "Fields", fname), "w") as f:
        f.write("\n".join("%.9g" % h for h in fields) + "\n")
    write_js(dirpath, keys["INTERACTIONS"])
    k = dict(keys)
    k.setdefault("TEMPNAME", tname)
    k.setdefault("FIELDNAME", fname)
    with open(os.path.join(dirpath, "INPUT.dat"), "w") as f:
        for key, v in k.items():
            f.write("%s %s\n" % (key, v))
    return os.path.join(dirpath, "INPUT.dat")


def write_config(dirpath, name, **overrides):
    c = CONFIGS[name]
    keys = dict(c["keys"])
    keys.update(overrides)
    return write_case(dirpath, keys, c["temps"], c["fields"])


if __name__ == "__main__":
    import sys
    out = sys.argv[2] if len(sys.argv) > 2 else "."
    print(write_config(out, sys.argv[1]))
