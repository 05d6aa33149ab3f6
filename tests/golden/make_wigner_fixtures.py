"""Generate Wigner 3j/6j/9j reference values with sympy (exact Racah algebra, an independent
library) for the oracle and the product to be checked against.  Run in a container that has
sympy; the JSON it writes is committed next to this script.

    python tests/golden/make_wigner_fixtures.py
"""
import itertools
import json
import random
from pathlib import Path

from sympy import Rational
from sympy.physics.wigner import wigner_3j, wigner_6j, wigner_9j

OUT = Path(__file__).with_name("wigner_fixtures.json")


def h(tj):
    return Rational(tj, 2)


def main():
    rnd = random.Random(20241202)
    six = []
    # exhaustive for 2j <= 4, then a seeded sample up to 2j = 12
    for args in itertools.product(range(0, 5), repeat=6):
        try:
            v = wigner_6j(*[h(x) for x in args])
        except ValueError:
            continue
        if v != 0:
            six.append([list(args), float(v.evalf(30))])
    while len(six) < 2500:
        args = [rnd.randint(0, 12) for _ in range(6)]
        try:
            v = wigner_6j(*[h(x) for x in args])
        except ValueError:
            continue
        if v != 0:
            six.append([args, float(v.evalf(30))])
    three = []
    while len(three) < 1500:
        tj = [rnd.randint(0, 10) for _ in range(3)]
        tm = [rnd.choice(range(-tj[i], tj[i] + 1, 2)) for i in range(2)]
        tm.append(-tm[0] - tm[1])
        if abs(tm[2]) > tj[2] or (tj[2] + tm[2]) % 2:
            continue
        try:
            v = wigner_3j(*[h(x) for x in tj], *[h(x) for x in tm])
        except ValueError:
            continue
        if v != 0:
            three.append([tj + tm, float(v.evalf(30))])
    nine = []
    while len(nine) < 300:
        args = [rnd.randint(0, 6) for _ in range(9)]
        try:
            v = wigner_9j(*[h(x) for x in args], prec=None)
        except (ValueError, TypeError):
            continue
        if v != 0:
            nine.append([args, float(v.evalf(30))])
    OUT.write_text(json.dumps({"generator": "sympy.physics.wigner", "6j": six, "3j": three, "9j": nine}))
    print(len(six), len(three), len(nine))


if __name__ == "__main__":
    main()
