"""CPU oracle for the X-PSI pulse-profile likelihood path.

Test infrastructure only: the product package never imports this module.
"""
