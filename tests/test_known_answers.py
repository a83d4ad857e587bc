"""The CPU oracle against its committed known answers
(tests/golden/oracle_known_answers.npz, written by
tests/golden/make_known_answers.py from the pinned oracle): every output bit
for bit, with inputs regenerated and checked against the stored ones."""
import os

import numpy as np

from tests.golden import make_known_answers as K

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "oracle_known_answers.npz")


def test_oracle_reproduces_known_answers(orc):
    fx = np.load(GOLD)
    c = K.cases()
    for k, v in c.items():
        assert np.array_equal(fx[k], v), k
    got = K.answers(orc)
    assert set(got) | set(c) == set(fx.files)
    for k, v in got.items():
        assert np.array_equal(fx[k], v), k
