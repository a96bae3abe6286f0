"""CPU oracle for the FBPINN train-step path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (`fbpinns_amd`) may
import, call or execute anything under `oracle/`.  The only permitted users are
`tests/`, `__graft_entry__.smoke()` (to check one step against it) and the `cpu_baseline` leg
of `bench.py` (as the timed CPU restatement of the method).

Modules
-------
takes    numpy restatement of the point->subdomain take builder
         (decompositions.py:217-227, decompositions_base.py:23-82,
         trainers.py:332-419, 482-507, 544-571).  Integer work, bit-exact.
model    torch (CPU, float64 or float32) restatement of FBPINN_model /
         FBPINN_forward / FBPINN_loss with the reference's nested-jvp
         derivative tree (trainers.py:63-107, 113-177, 197-267).
problems restatement of the config problems' constraining_fn / loss_fn
         (problems.py:149-481) and the multiscale Poisson problem.
adam     numpy restatement of optax.adam's update (scale_by_adam +
         scale_by_learning_rate + apply_updates), used at trainers.py:294-295.
exact    exact / reference solutions used for L1 (problems.py:134-146,
         burgers_solution.py:40-144, seismic_CPML_2D_pressure_second_order.py).

How each function is pinned (checks that a plausible mistake would fail) is
listed in the header of each module and in tests/test_oracle_*.py.
"""
