"""Progress lines and measurement records of the long GPU tests.

Progress lines go to pytest's terminal writer, which prints at once even while the test's own
output is captured, so a test that runs for minutes keeps reporting.  Records (the errors and
ratios a test measured) are written as JSON into the directory named by the environment
variable KMI_TEST_RECORDS, when it is set."""
import json
import os

_config = None


def set_writer(config):
    """Keep pytest's config; its terminal reporter is looked up when a line is written (it may
    register after the conftest that calls this)."""
    global _config
    _config = config


def progress(msg):
    print(msg, flush=True)
    writer = _config.pluginmanager.get_plugin("terminalreporter") if _config is not None else None
    if writer is not None:
        writer.write_line(msg)
    d = os.environ.get("KMI_TEST_RECORDS")
    if d:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "progress.log"), "a") as fh:
            fh.write(msg + "\n")


def record(filename, name, data):
    """Merge data under `name` into the JSON file `filename` of the records directory."""
    d = os.environ.get("KMI_TEST_RECORDS")
    if not d:
        return
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, filename)
    cur = json.load(open(path)) if os.path.exists(path) else {}
    cur[name] = data
    with open(path, "w") as fh:
        json.dump(cur, fh, indent=1)
