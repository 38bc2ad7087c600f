from .fd2d import fd2d

__all__ = ['fd2d']
