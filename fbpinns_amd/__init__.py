"""fbpinns_amd: an MI355X-native FBPINN training engine.

Keeps the reference's Problem / Decomposition / Network / Domain /
ActiveScheduler / Constants / FBPINNTrainer API; the train step runs as
hand-written HIP kernels (libfbpx.so, C ABI in include/fbpx.h).
"""

__version__ = "0.1.0"
