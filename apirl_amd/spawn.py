"""Running other programs from a process that uses the GPU.

A process that has initialised the GPU must not fork and exec another program (the
child inherits the device state; on MI355X hosts such an exec is refused or brings the
process down).  Launcher.start() therefore creates, before the GPU is touched, a
helper process from a fresh interpreter (multiprocessing "forkserver"); run() later
asks that helper to start the program and returns its exit code and output."""
import multiprocessing as _mp
import subprocess as _sp


def _run(args, env, timeout, cwd):
    p = _sp.run(args, env=env, capture_output=True, text=True, timeout=timeout, cwd=cwd)
    return p.returncode, p.stdout, p.stderr


class Launcher:
    def __init__(self):
        self._pool = None

    def start(self):
        """Starts the helper; call before anything initialises the GPU."""
        if self._pool is None:
            ctx = _mp.get_context("forkserver")
            ctx.set_forkserver_preload([])
            self._pool = ctx.Pool(1)
        return self

    @property
    def started(self):
        return self._pool is not None

    def run(self, args, env=None, timeout=600, cwd=None):
        """(returncode, stdout, stderr) of `args`, started by the helper."""
        if self._pool is None:
            raise RuntimeError("Launcher.start() must run before the GPU is initialised")
        return self._pool.apply(_run, (list(args), env, timeout, cwd))

    def close(self):
        if self._pool is not None:
            self._pool.close()
            self._pool.join()
            self._pool = None
