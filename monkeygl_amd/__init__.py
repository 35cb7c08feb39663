"""MI355X-native render path of MonkeyGL: HIP kernels behind a C ABI.

``monkeygl_amd._capi`` binds ``include/monkeygl_hip.h``; ``params`` restates the
reference's kernel-argument producers; ``workloads`` builds seeded synthetic
volumes.
"""
