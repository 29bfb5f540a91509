# __init__.py for residuals package
