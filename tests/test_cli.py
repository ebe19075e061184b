"""The command-line driver (superman_amd/superman_gpu): argument and file
handling of the reference's CLI (IO/IO.h:163-596), checked without a GPU."""
import numpy as np

import cli


def test_unknown_algorithm_and_mode():
    rc, out, err = cli.run(np.eye(3), "algorithm=fastest")
    assert rc == 1 and "UNKNOWN ALGORITHM: fastest" in err
    rc, out, err = cli.run(np.eye(3), "mode=quantum")
    assert rc == 1 and "UNKNOWN MODE: quantum" in err


def test_missing_file_and_filename():
    rc, out, err = cli.run(None, path="/nonexistent/m.mtx")
    assert rc == 1 and "File could not be opened" in err
    import subprocess
    p = subprocess.run([cli.binary()], capture_output=True, text=True)
    assert p.returncode == 1 and "absolute path of the matrix file" in p.stderr


def test_settings_read_as_the_reference_reads_them():
    # unknown keys are reported and skipped (IO.h:382-386); bad integers and
    # precisions end the run
    rc, out, err = cli.run(np.eye(3), "colour=blue", "scaling=true", "scale_into=4", "mode=cpu")
    assert "UNKNOWN ARGUMENT: colour - skipping it." in out and "CPU mode" in err
    rc, out, err = cli.run(np.eye(3), "thread_count=many")
    assert rc == 1 and "integer value should be provided to the thread_count" in err
    rc, out, err = cli.run(np.eye(3), "calculation_precision=quad")
    assert rc == 1 and "Calculation precision can either be kahan or dd" in err
    rc, out, err = cli.run(np.eye(3), "calculation_precision=256", "device_id=99")
    assert rc == 1 and "Permanent:" not in out


def test_cpu_mode_is_not_served():
    rc, out, err = cli.run(np.eye(4), "mode=cpu")
    assert rc == 1 and "CPU mode" in err


def test_missing_device_is_an_error_not_a_fallback():
    rc, out, err = cli.run(np.ones((6, 6)), "device_id=99")
    assert rc == 1 and "Permanent:" not in out
