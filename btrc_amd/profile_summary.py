"""Turns rocprofv3 result databases into the short text summaries kept in
profiles/.

    python btrc_amd/profile_summary.py kernels <run_results.db> [title] [command]
    python btrc_amd/profile_summary.py traffic <fetch.db> <write.db> <copy_fetch.db> <copy_write.db>
                                       <rays_per_dispatch> <kernel substring> [command]
    python btrc_amd/profile_summary.py frame_traffic <fetch.db> <write.db> <copy_fetch.db> <copy_write.db>
                                       <total_rays> <kernel substring> [command]

`kernels` prints per-kernel calls, total and average duration.
`traffic` scales FETCH_SIZE / WRITE_SIZE (KiB per dispatch) by what the same
counters read for a device-to-device copy of known size, and prints bytes per
ray of the named kernel as `hbm_bytes_per_ray: <value>`.
"""
import sqlite3
import sys


def kernels(db, title="kernels", command=None):
    con = sqlite3.connect(db)
    rows = list(con.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    out = [f"# {title}", ""]
    if command:
        out += ["Command:", "", "    " + command, ""]
    out += ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for name, calls, total, avg, pct in rows:
        short = name.replace("btrc::(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        # durations in the database are in microseconds here
        out.append(f"| {short} | {calls} | {total / 1e3:.3f} | {avg:.1f} | {pct:.1f} |")
    return "\n".join(out)


def counter_values(db, kernel_substring):
    con = sqlite3.connect(db)
    return [v for (n, v) in con.execute("select kernel_name, value from counters_collection")
            if kernel_substring in n]


def traffic(fetch_db, write_db, copy_fetch_db, copy_write_db, rays, kernel, command=None):
    copy_kib = 1024.0 * 1024.0  # 1 GiB copied: that many KiB read and written
    cf = max(counter_values(copy_fetch_db, "copyBuffer"))
    cw = max(counter_values(copy_write_db, "copyBuffer"))
    f_scale, w_scale = copy_kib / cf, copy_kib / cw
    f = counter_values(fetch_db, kernel)
    w = counter_values(write_db, kernel)
    fb = sum(f) / len(f) * f_scale * 1024.0
    wb = sum(w) / len(w) * w_scale * 1024.0
    per_ray = (fb + wb) / rays
    out = ["# HBM traffic of the extension-ray traversal kernel", ""]
    if command:
        out += ["Commands:", ""] + ["    " + c for c in command.split(" ;; ")] + [""]
    out += [f"- copy calibration: 1 GiB device copy reads FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB "
            f"-> scale factors fetch x{f_scale:.4f}, write x{w_scale:.4f}",
            f"- kernel `{kernel}`: {len(f)} dispatches of {rays} rays, FETCH_SIZE {sum(f) / len(f):.1f} KiB, "
            f"WRITE_SIZE {sum(w) / len(w):.1f} KiB per dispatch",
            f"- fetched bytes per ray: {fb / rays:.1f}",
            f"- written bytes per ray: {wb / rays:.1f}",
            f"hbm_bytes_per_ray: {per_ray:.1f}"]
    return "\n".join(out)


def frame_traffic(fetch_db, write_db, copy_fetch_db, copy_write_db, total_rays, kernel, command=None):
    """bytes per ray over all dispatches of `kernel` in a rendered frame"""
    copy_kib = 1024.0 * 1024.0
    cf = max(counter_values(copy_fetch_db, "copyBuffer"))
    cw = max(counter_values(copy_write_db, "copyBuffer"))
    f_scale, w_scale = copy_kib / cf, copy_kib / cw
    f = counter_values(fetch_db, kernel)
    w = counter_values(write_db, kernel)
    fb = sum(f) * f_scale * 1024.0
    wb = sum(w) * w_scale * 1024.0
    out = [f"# HBM traffic of {kernel} over one rendered frame", ""]
    if command:
        out += ["Commands:", ""] + ["    " + c for c in command.split(" ;; ")] + [""]
    out += [f"- copy calibration: 1 GiB device copy reads FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB "
            f"-> scale factors fetch x{f_scale:.4f}, write x{w_scale:.4f}",
            f"- kernel `{kernel}`: {len(f)} dispatches, {total_rays} rays in total, FETCH_SIZE {sum(f):.1f} KiB, "
            f"WRITE_SIZE {sum(w):.1f} KiB summed over the dispatches",
            f"- fetched bytes per ray: {fb / total_rays:.1f}",
            f"- written bytes per ray: {wb / total_rays:.1f}",
            f"hbm_bytes_per_ray: {(fb + wb) / total_rays:.1f}"]
    return "\n".join(out)


def main(argv):
    if argv[1] == "kernels":
        print(kernels(argv[2], argv[3] if len(argv) > 3 else "kernels", argv[4] if len(argv) > 4 else None))
    elif argv[1] == "traffic":
        print(traffic(argv[2], argv[3], argv[4], argv[5], int(argv[6]), argv[7],
                      argv[8] if len(argv) > 8 else None))
    elif argv[1] == "frame_traffic":
        print(frame_traffic(argv[2], argv[3], argv[4], argv[5], int(argv[6]), argv[7],
                            argv[8] if len(argv) > 8 else None))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv)
