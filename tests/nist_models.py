"""The eight NIST StRD datasets with certified fits (data-sources/NIST/*/*.dat of the
reference): each model's formula (FP64, numpy) and, where it is not the rational
function of the differentiation workload, its postfix tape over the config grammar with
the certified parameters as float32 constants."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def C(v):
    return ("CONST", float(np.float32(v)))


MODELS = {
    "BoxBOD": (lambda b, x: b[0] * (1 - np.exp(-b[1] * x)),
               lambda b: [C(b[0]), "ONE", C(-b[1]), "X", "MUL", "EXP", "SUB", "MUL"]),
    "MGH09": (lambda b, x: b[0] * (x ** 2 + x * b[1]) / (x ** 2 + x * b[2] + b[3]),
              lambda b: [C(b[0]), "X", "X", "MUL", "X", C(b[1]), "MUL", "ADD", "MUL", "X", "X", "MUL", "X", C(b[2]),
                         "MUL", "ADD", C(b[3]), "ADD", "DIV"]),
    "MGH10": (lambda b, x: b[0] * np.exp(b[1] / (x + b[2])),
              lambda b: [C(b[0]), C(b[1]), "X", C(b[2]), "ADD", "DIV", "EXP", "MUL"]),
    "Eckerle4": (lambda b, x: (b[0] / b[1]) * np.exp(-0.5 * ((x - b[2]) / b[1]) ** 2),
                 lambda b: [C(b[0]), C(b[1]), "DIV", C(-0.5), "X", C(b[2]), "SUB", C(b[1]), "DIV", "X", C(b[2]), "SUB",
                            C(b[1]), "DIV", "MUL", "MUL", "EXP", "MUL"]),
    "Bennett5": (lambda b, x: b[0] * (b[1] + x) ** (-1 / b[2]),
                 lambda b: [C(b[0]), C(b[1]), "X", "ADD", C(-1 / b[2]), "POW", "MUL"]),
    "Rat42": (lambda b, x: b[0] / (1 + np.exp(b[1] - b[2] * x)),
              lambda b: [C(b[0]), "ONE", C(b[1]), C(b[2]), "X", "MUL", "SUB", "EXP", "ADD", "DIV"]),
    "Rat43": (lambda b, x: b[0] / ((1 + np.exp(b[1] - b[2] * x)) ** (1 / b[3])),
              lambda b: [C(b[0]), "ONE", C(b[1]), C(b[2]), "X", "MUL", "SUB", "EXP", "ADD", C(1 / b[3]), "POW", "DIV"]),
    "Thurber": (lambda b, x: (b[0] + b[1] * x + b[2] * x ** 2 + b[3] * x ** 3) / (1 + b[4] * x + b[5] * x ** 2 + b[6] * x ** 3),
                None),
}


def certified(name):
    """(record of tests/golden/nist_certified.json, parameter list b1..bk)."""
    cert = json.load(open(os.path.join(GOLDEN, "nist_certified.json")))[name]
    return cert, [cert["params"][f"b{i + 1}"] for i in range(len(cert["params"]))]


def tape(name, b):
    """Postfix tape (ops u8, cs f32) of the certified model with parameters b."""
    import oracle_lib as O
    f, t = MODELS[name]
    if t is None:
        from gpuropolis_amd.workloads import rational_tape
        return rational_tape(np.array(b))
    return O.tape_arrays(t(b))
