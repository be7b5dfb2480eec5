"""CPU restatement oracle of river-basin-sim's solver time step.

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import it; the product package never does.
"""
