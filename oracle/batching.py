"""Stack batching helpers (oracle restatement).

Follows ``src/qttools/utils/mpi_utils.py:16-67`` (``get_section_sizes``) and
``src/qttools/utils/solvers_utils.py:7-30`` (``get_batches``).
"""

import numpy as np


def get_section_sizes(num_elements: int, num_sections: int, strategy: str = "balanced"):
    """Split ``num_elements`` into ``num_sections`` parts (mpi_utils.py:16-67)."""
    quotient, remainder = divmod(num_elements, num_sections)
    if strategy == "balanced":
        sizes = [quotient + 1] * remainder + [quotient] * (num_sections - remainder)
    elif strategy == "greedy":
        sizes = []
        for _ in range(num_sections):
            sizes.append(min(quotient + min(remainder, 1), num_elements - sum(sizes)))
    else:
        raise ValueError(f"Invalid strategy: {strategy}")
    return sizes, max(sizes) * num_sections


def get_batches(num_sections: int, max_batch_size: int):
    """Batch sizes and offsets over the first stack axis (solvers_utils.py:7-30).

    The number of batches is ``num_sections // min(max_batch_size,
    num_sections)`` with balanced sizes, so a batch may exceed
    ``max_batch_size`` (e.g. 150 with max 100 gives one batch of 150).
    """
    sizes, _ = get_section_sizes(num_sections, num_sections // min(max_batch_size, num_sections))
    slices = np.hstack(([0], np.cumsum(np.array(sizes))))
    return sizes, slices
