"""End-of-run report comparison (results.yaml, results.txt, stdout).

The reports of the reference (ref/profile.c:37-1300) and of the drop-in
print the same lines; the numbers that measure time or memory differ by
nature (host wall time against GPU device time, mallocs against device
buffers), every other number -- configuration, object motion, faces
exchanged, refinement steps, blocks, floating-point operations -- must be
equal.  mask_text() and mask_yaml() keep a report's structure and those
numbers and replace the measured ones with "T".
"""
import os
import re
import subprocess
import tempfile

NUM = re.compile(r"-?(?:\d+\.\d+|inf|nan)")

# text report: labels whose numbers are timings, rates or memory
TEXT_MEASURED = re.compile(
    r"^\s*(Total time for test|Number of malloc calls|Amount malloced|Malloc calls in|"
    r"Main \(parse, allocate\) Time|Initailize Time|Time:|total GFLOPS|Average GFLOPS|"
    r"Sum of min/max|red :|calc:|Total\s+:|Post IRecv|Pack faces|Send messages|"
    r"Exchange same level|Exchange diff level|Apply BC|Wait time|Unpack faces|"
    r"initial refine|compare objects|mark refine/coarsen|communicate block|split blocks|"
    r"sync time|time for groups|misc time|total coarsen blocks|coarsen blocks|pack blocks|"
    r"move blocks|unpack blocks|total redistribute|choose blocks|total load balance|sort |"
    r"move dots back|misc )")
SUMMARY = re.compile(r"^(Summary: ranks \d+ ts \d+) time \S+ calc \S+ comm \S+ red \S+ "
                     r"refine \S+ (blocks/ts \S+ max_blocks \d+)")

# yaml report: keys whose values are timings, rates or memory
YAML_MEASURED = re.compile(
    r"^\s*(total_time|memory_used|compute_time|total_gflops|ave_gflops|"
    r"(total|x|y|z)_comm_(ave|min|max)|\w*post_recv|\w*pack_faces|\w*send_mess|"
    r"(total|x|y|z)_exch_(same|diff)|\w*apply_bc|\w*wait_time|\w*unpack_faces|gridsum|"
    r"refine_time|time_|plot_time)")


def mask_text(text):
    out = []
    for line in text.splitlines():
        m = SUMMARY.match(line)
        if m:
            out.append("%s time T calc T comm T red T refine T %s" % (m.group(1), m.group(2)))
        elif TEXT_MEASURED.match(line):
            out.append(NUM.sub("T", line))
        else:
            out.append(line)
    return out


def mask_yaml(text):
    out = []
    for line in text.splitlines():
        if YAML_MEASURED.match(line):
            out.append(line.split(":")[0] + ": T")
        else:
            out.append(line)
    return out


def report_part(stdout):
    """The report lines of a run's stdout: from the report's start to the
    Summary line (per-timestep lines before it are compared elsewhere)."""
    lines = stdout.splitlines()
    start = next(i for i, l in enumerate(lines) if "Start report" in l)
    end = next(i for i, l in enumerate(lines) if l.startswith("Summary:"))
    return "\n".join(lines[start - 1:end + 1])


def run_reports(exe, args, env=None):
    """Run `exe args --report_perf 15` in a scratch directory; return the
    masked stdout report, results.txt and results.yaml as line lists, and
    the lines printed before the report (per-timestep block counts)."""
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run([exe] + args.split() + ["--report_perf", "15"], cwd=d,
                           capture_output=True, text=True, timeout=600, env=env)
        if p.returncode:
            raise RuntimeError("%s failed (%d): %s" % (exe, p.returncode, p.stderr[-2000:]))
        with open(os.path.join(d, "results.txt")) as f:
            txt = f.read()
        with open(os.path.join(d, "results.yaml")) as f:
            yml = f.read()
    report = report_part(p.stdout)
    before = p.stdout.split("\n ================ Start report")[0].splitlines()
    return {"stdout": mask_text(report), "txt": mask_text(txt), "yaml": mask_yaml(yml),
            "progress": before}
