"""Shared test cases (data only)."""


def empty_spin_case(nO, case):
    """(occupations, nC, nR) per spin of a case with one spin without active
    pairs, for closed-shell orbitals with nO doubly occupied."""
    return {"beta empty": ((nO + 1, 0), (0, 0), (0, 0)),
            "beta empty, frozen core and virtual": ((nO, 0), (1, 0), (0, 1)),
            "beta occupied all frozen": ((nO, nO), (0, nO), (0, 0)),
            "alpha occupied all frozen": ((nO, nO), (nO, 0), (0, 0))}[case]


EMPTY_SPIN = ["beta empty", "beta empty, frozen core and virtual", "beta occupied all frozen",
              "alpha occupied all frozen"]
