"""Small host-side helpers shared by the solvers."""

import numpy as np


def get_section_sizes(num_elements: int, num_sections: int, strategy: str = "balanced"):
    """Split ``num_elements`` into ``num_sections`` parts.

    Same contract as src/qttools/utils/mpi_utils.py:16-67 (``balanced``:
    the first ``remainder`` sections get one extra element; ``greedy``: as
    many full sections as possible).  Returns (sizes, max(sizes) * sections).
    """
    q, r = divmod(num_elements, num_sections)
    if strategy == "balanced":
        sizes = [q + 1] * r + [q] * (num_sections - r)
    elif strategy == "greedy":
        sizes = []
        for _ in range(num_sections):
            sizes.append(min(q + min(r, 1), num_elements - sum(sizes)))
    else:
        raise ValueError(f"Invalid strategy: {strategy}")
    return sizes, max(sizes) * num_sections


def get_batches(num_sections: int, max_batch_size: int):
    """Batches over the first stack axis (src/qttools/utils/solvers_utils.py:7-30).

    ``num_sections // min(max_batch_size, num_sections)`` balanced batches, so
    a batch can exceed ``max_batch_size`` (150 with max 100 -> one batch).
    """
    sizes, _ = get_section_sizes(num_sections, num_sections // min(max_batch_size, num_sections))
    return sizes, np.hstack(([0], np.cumsum(np.array(sizes, dtype=np.int64))))
