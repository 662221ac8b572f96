"""Plain-text log with the summary-table interface the ABFS classes print through.

Mirrors the part of pCore's LogFileWriter used on this path
(pCore-1.9.0/pCore/LogFileWriter.py:39-341): ``GetSummary().Start/Entry/Stop``
and ``LogFileActive``.
"""

import sys


class _Summary:
    def __init__(self, log, pageWidth=80, valueWidth=14):
        self.log = log
        self.entries = []
        self.title = ""
        self.width = pageWidth
        self.valueWidth = valueWidth

    def Start(self, title):
        self.title = title
        self.entries = []

    def Entry(self, label, value):
        self.entries.append((str(label), str(value)))

    def Stop(self):
        w = self.width
        out = ["-" * w, self.title.center(w), "-" * w]
        half = w // 2
        row = ""
        for k, (label, value) in enumerate(self.entries):
            cell = "{:<{lw}}{:>{vw}}".format(label, value, lw=half - self.valueWidth - 2, vw=self.valueWidth)
            row = cell if k % 2 == 0 else row + "  " + cell
            if k % 2 == 1:
                out.append(row)
                row = ""
        if row:
            out.append(row)
        out.append("-" * w)
        self.log.Text("\n".join(out) + "\n")


class TextLogFile:
    """Writes summaries to a stream (stdout by default)."""

    def __init__(self, stream=None, active=True):
        self.stream = stream if stream is not None else sys.stdout
        self.active = active
        self.lines = []

    def GetSummary(self, pageWidth=80, valueWidth=14):
        return _Summary(self, pageWidth, valueWidth)

    def Text(self, text):
        self.lines.append(text)
        if self.stream is not None:
            self.stream.write(text)

    def Paragraph(self, text):
        self.Text(text + "\n")


logFile = TextLogFile(active=True)


def LogFileActive(log):
    return log is not None and getattr(log, "active", False)
