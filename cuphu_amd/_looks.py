"""Effective-looks computation (reference: src/cuphu/_looks.py:83-88)."""

__all__ = ["get_effective_looks"]


def get_effective_looks(
    range_looks: float,
    azimuth_looks: float,
    range_spacing: float,
    azimuth_spacing: float,
    range_resolution: float,
    azimuth_resolution: float,
) -> float:
    r"""Equivalent number of independent looks for ``unwrap``'s ``nlooks``:

    .. math:: n_e = k_r k_a \frac{d_r d_a}{\rho_r \rho_a}

    >>> get_effective_looks(13, 16, 2.33, 2.75, 7.5, 6.0)
    29.616888888888894
    """
    return (range_looks * azimuth_looks * (range_spacing * azimuth_spacing)
            / (range_resolution * azimuth_resolution))
